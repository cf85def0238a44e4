/*
 * wave_tables.h -- topology tables of the one-world-per-wavefront kernel,
 * derived on the host from a built world: joint dependency levels of the
 * Gauss-Seidel sweep, spheres grouped per body, contact-list slots.
 */
#pragma once
#include "lane_engine.h"

namespace xpbd {

constexpr int WAVE_MAX_BODIES = 16, WAVE_MAX_JOINTS = 16, WAVE_MAX_SPHERES = 48, WAVE_MAX_SLOTS = 128;

struct WaveK {
    int nb, nj, ns, ncb;
    int jp[WAVE_MAX_JOINTS], jc[WAVE_MAX_JOINTS];  /* parent (-1: static world), child */
    int nlev;
    int lev_first[WAVE_MAX_JOINTS + 1];            /* joints of level L: lev_joint[lev_first[L] .. lev_first[L+1]) */
    int lev_joint[WAVE_MAX_JOINTS];
    int cb_body[WAVE_MAX_BODIES];                  /* body of contact group g */
    int cb_first[WAVE_MAX_BODIES + 1];             /* its spheres: cb_sphere[cb_first[g] .. cb_first[g+1]) */
    int cb_sphere[WAVE_MAX_SPHERES];
    int cb_slot[WAVE_MAX_BODIES], cb_cap[WAVE_MAX_BODIES];
    int s_group[WAVE_MAX_SPHERES];                 /* contact group of each sphere */
    int serial_collide;                            /* the 128-record cap could bind: keep sphere order */
};

struct HostWorld;
/* false if the world exceeds the kernel's capacities */
bool build_wave_tables(const HostWorld &w, float cell, WaveK &T);

struct DevWorlds;
struct Model;
hipError_t launch_step_wave(const DevWorlds &d, const Model *m_dev, const WaveK *wk_dev, const double *targets_d,
                            const float *targets_f, int grid_blocks, hipStream_t s);

} /* namespace xpbd */
