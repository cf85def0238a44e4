/*
 * sim_internal.h -- the arc_sim handle shared by the host sources of
 * libarcsim (sim.cpp: the arc_sim.h ABI; batch.cpp: batched, persistent and
 * graph launches and the autotuner; server.cpp: the step server).
 */
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/arc_gpu.h"
#include "device_state.h"
#include "setup.h"
#include "wave_tables.h"

namespace xpbd {
/* whether qdiv's three-instruction quotient by d equals the IEEE quotient for
 * every binary32 numerator of its domain on the current device: an exhaustive
 * device check (selfcheck.hip), run once per device and divisor */
bool quotient_checked(float d);
} /* namespace xpbd */

using namespace xpbd;

struct BatchState;

struct arc_sim {
    int n;
    int substeps;
    double dt_d, cell_d;
    SimSetup su;
    DevWorlds dev;
    Model *d_model;
    ResetK *d_reset;
    uint32_t *d_jtab;      /* the reset's jump tables (ResetK::jtab of d_reset) */
    TemplateK *d_tmpl;
    hipStream_t stream;
    int device;
    /* staging */
    double *d_io;          /* device: max(N*12, N*64, N*k) doubles */
    size_t d_io_doubles;
    char *d_aux;           /* device: ids/seeds/difficulty/offsets */
    size_t d_aux_bytes;
    char *h_pin;           /* pinned host staging */
    size_t h_pin_bytes;
    hipEvent_t up_done;    /* arc_sim_step's upload out of h_pin finished */
    bool up_pending;       /* up_done recorded and not yet waited for */
    float *d_dump;
    arc_world *mirror;
    int block, grid_blocks;
    int tile;              /* LDS terrain tile side of the lane form: 0 or 16 */
    int persistent;        /* lane form: wavefronts take work from a counter */
    int form;              /* 0: one world per lane, 1: one world per wavefront, 2: four lanes per world
                              (set at create), 3: two lanes per world, 4: four lanes per world with the
                              hip joints of a block's worlds solved through LDS */
    WaveK wave;
    WaveK *d_wave;
    TaskK task;
    TaskState ts;
    void *ts_mem;
    int reset_blocks;
    int err;
    char msg[256];
    BatchState *batch;     /* batched / graph form (batch.cpp), created on first use */
    int contact_layout;    /* where the last step left its contact lists: 0 nowhere readable (the
                              allocation-free step), 1 cbuf (one world per lane, plain form), 2 lcbuf with
                              four lanes per world, 3 lcbuf with two */
};

/* records the contact-list layout of the configured form, before its step is enqueued */
inline void note_contacts(arc_sim *s) { s->contact_layout = s->form <= 1 ? 1 : s->form == 3 ? 3 : 2; }

/* records the first error of a handle (sticky) and prints it */
void set_error(const arc_sim *cs, int code, const char *what, hipError_t e);
/* the configured form of the step kernel over all worlds, float targets */
hipError_t launch_form(arc_sim *s, const float *targets);
/* releases the batch state of a handle */
void batch_free(arc_sim *s);
/* every world of the sim in the configured register-resident form (lane, leg,
 * pair or shared) and distribution */
hipError_t launch_lane_all(arc_sim *s, const float *targets);
/* the leg form's contact lists (allocated on first use) */
hipError_t ensure_leg(arc_sim *s);
/* [p, p + bytes) lies inside one allocation of the sim's GPU (device or
 * managed memory); bytes = 0 checks the address only */
bool device_memory(const arc_sim *s, const void *p, size_t bytes);
/* arcx_step_batch_rows for rows the caller has already checked with
 * device_memory (the step server checks them when they are submitted) */
int step_batch_rows_checked(arc_sim *s, const int *starts, const int *counts, int n_req, const float *const *rows,
                            int task, int graph);
/* buffers the configured launch allocates on first use, made before a capture */
hipError_t prepare_launch(arc_sim *s);

#define HIPCHK(s, call, what)                                 \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) {                               \
            set_error((s), (int)e_, (what), e_);              \
            return;                                           \
        }                                                     \
    } while (0)
#define HIPCHK_RET(s, call, what, ret)                        \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) {                               \
            set_error((s), (int)e_, (what), e_);              \
            return (ret);                                     \
        }                                                     \
    } while (0)
