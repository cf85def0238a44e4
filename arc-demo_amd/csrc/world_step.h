/*
 * world_step.h -- one hand-built world (any bodies, hinges and spheres within
 * the wavefront kernel's tables) stepped on the device by the plain form's
 * substep code (wave_kernel.hip).  The host describes the world in a
 * HostWorld (model.h) and owns it between steps; a step uploads the dynamic
 * state, runs `steps` world steps in LDS and downloads it again.
 */
#pragma once
#include "wave_tables.h"

namespace xpbd {

/* constants of one world for one substep length */
struct WorldModel {
    BodyK body[WAVE_MAX_BODIES];
    JointK joint[WAVE_MAX_JOINTS];
    SphereK sphere[WAVE_MAX_SPHERES];
    TerrainK terr;
    V3 g;
    float mus, mud;
    float dt, h, hh, inv_h, two_h;
    int substeps;
    float outside; /* height outside the grid */
};

/* dynamic state of the world, exchanged with the host around a step */
struct WorldState {
    Body b[WAVE_MAX_BODIES]; /* x, q, v, w (the caches are rebuilt by each step) */
    float target[WAVE_MAX_JOINTS], angle[WAVE_MAX_JOINTS], rate[WAVE_MAX_JOINTS], torque[WAVE_MAX_JOINTS];
    int touch[WAVE_MAX_SPHERES], anchored[WAVE_MAX_SPHERES];
    V3 anc[WAVE_MAX_SPHERES], ancn[WAVE_MAX_SPHERES];
    int unstable;
    int overflow; /* a contact group outgrew its list (never, by the tables' sizing) */
    float time;
};

/* Runs `steps` world steps (csrc/arc_phys.c:707-747) of the world in *st on
 * one wavefront; heights: the terrain's nx * ny columns, row-major. */
hipError_t launch_world_step(WorldState *st, const WorldModel *m, const WaveK *wk, const float *heights, int steps,
                             hipStream_t s);

} /* namespace xpbd */
