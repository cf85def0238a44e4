/*
 * setup.cpp -- per-simulation constants (see setup.h).
 */
#include "setup.h"
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>

namespace xpbd {

/* The device engine hard-codes the quadruped's topology; check the built
 * robot against it before trusting the constants. */
static bool topology_ok(const HostWorld &w)
{
    if (w.nb != NB || w.nj != NJ || w.ns != NS)
        return false;
    for (int b = 0; b < NB; ++b) {
        bool iso = body_kind(b) == KISO;
        if ((w.bk[b].iso_I != 0) != iso || (w.bk[b].iso_invI != 0) != iso)
            return false;
    }
    for (int j = 0; j < NJ; ++j)
        if (w.j[j].parent != joint_parent(j) || w.j[j].child != joint_child(j))
            return false;
    for (int s = 0; s < NS; ++s) {
        const HostSphere &sp = w.s[s];
        if (sp.body != sphere_body(s))
            return false;
        if (sphere_at(s) == AT_ZERO && (sp.at.x != 0.0f || sp.at.y != 0.0f || sp.at.z != 0.0f))
            return false;
        if (sphere_at(s) == AT_Z && (sp.at.x != 0.0f || sp.at.y != 0.0f))
            return false;
    }
    return true;
}

static bool same_bits(const void *a, const void *b, size_t n) { return memcmp(a, b, n) == 0; }

/* The step kernel compiles the robot's constants in (robot_consts.h); they must
 * be the bits of the model built here. */
static bool constants_match(const HostWorld &w)
{
    for (int b = 0; b < NB; ++b)
        if (!same_bits(&robot::body[b], &w.bk[b], sizeof(BodyK)))
            return false;
    for (int j = 0; j < NJ; ++j) {
        JointK a = robot::joint[j], k = joint_constants(w.j[j], 1.0f);
        a.soft = k.soft = a.cap_pos = k.cap_pos = a.tmax_h = k.tmax_h = 0.0f;
        if (!same_bits(&a, &k, sizeof(JointK)))
            return false;
    }
    for (int s = 0; s < NS; ++s) {
        const SphereK &k = robot::sphere[s];
        if (!same_bits(&k.at, &w.s[s].at, sizeof(V3)) || !same_bits(&k.rad, &w.s[s].rad, sizeof(float))
            || k.body != w.s[s].body)
            return false;
    }
    return true;
}

bool sim_setup(SimSetup &S, double dt, int substeps, int grid, double cell)
{
    if (substeps <= 0 || grid <= 0 || !(cell > 0.0) || !(dt > 0.0)) {
        fprintf(stderr, "libarcsim: invalid simulation parameters\n");
        return false;
    }
    S.grid = grid < GRID_MAX ? grid : GRID_MAX;
    world_init(S.tmpl);
    robot_build(S.tmpl, S.info);
    terrain_flat(S.tmpl, S.grid, S.grid, (float)cell, 0.0f);
    if (!topology_ok(S.tmpl)) {
        fprintf(stderr, "libarcsim: robot topology does not match the device engine\n");
        return false;
    }
    if (!constants_match(S.tmpl)) {
        fprintf(stderr, "libarcsim: compiled-in robot constants differ from the built robot (rebuild)\n");
        return false;
    }
    for (int i = 0; i < S.tmpl.ns; ++i)
        S.channel[i] = i < S.info.n_torso_spheres ? ROBOT_LEGS : ROBOT_LEGS + 1;
    for (int i = 0; i < ROBOT_LEGS; ++i)
        S.channel[S.info.foot[i]] = i;
    for (int i = 0; i < S.tmpl.ns; ++i)
        if (S.channel[i] != sphere_channel(i)) {
            fprintf(stderr, "libarcsim: sphere channels do not match the device engine\n");
            return false;
        }

    Model &M = S.model;
    memset(&M, 0, sizeof(M));
    const float dtf = (float)dt;
    const float h = dtf / (float)substeps;
    for (int b = 0; b < NB; ++b)
        M.body[b] = S.tmpl.bk[b];
    for (int j = 0; j < NJ; ++j)
        M.joint[j] = joint_constants(S.tmpl.j[j], h);
    for (int k = 0; k < NS; ++k) {
        M.sphere[k].at = S.tmpl.s[k].at;
        M.sphere[k].rad = S.tmpl.s[k].rad;
        M.sphere[k].body = S.tmpl.s[k].body;
    }
    M.terr = S.tmpl.t;
    M.g = S.tmpl.g;
    M.mus = S.tmpl.mus;
    M.mud = S.tmpl.mud;
    M.dt = dtf;
    M.h = h;
    M.hh = h * h;
    M.inv_h = 1.0f / h;
    M.two_h = 2.0f / h;
    M.inv_sub = 1.0f / (float)substeps;
    M.substeps = substeps;
    int slot = 0;
    for (int ci = 0; ci < NCB; ++ci) {
        int cap = 0;
        for (int k = 0; k < NS; ++k)
            if (S.tmpl.s[k].body == cbody(ci))
                cap += window_cells(S.tmpl.s[k].rad, (float)cell);
        cap = std::min(cap, MAXC);
        M.cslot[ci] = slot;
        M.ccap[ci] = cap;
        slot += cap;
    }
    M.cslots = slot;

    ResetK &R = S.rk;
    for (int j = 0; j < NJ; ++j) {
        R.def[j] = S.info.def[j];
        R.lo[j] = S.info.lo[j];
        R.hi[j] = S.info.hi[j];
    }
    R.stand = S.info.stand;
    R.footprint = S.info.footprint;
    R.grid = S.grid;
    R.cell = (float)cell;
    /* jump tables of a generated reset's stream cut into RESET_LANES chunks */
    R.chunk = reset_chunk(S.grid);
    S.jtab.assign((size_t)RESET_LANES * JTAB_WORDS, 0u);
    rng_lane_tables(R.chunk, S.jtab.data());
    R.jtab = S.jtab.data();

    TemplateK &T = S.tk;
    for (int b = 0; b < NB; ++b) {
        T.x[b] = S.tmpl.b[b].x;
        T.q[b] = S.tmpl.b[b].q;
    }
    for (int j = 0; j < NJ; ++j) {
        T.target[j] = S.tmpl.j[j].target;
        T.angle[j] = S.tmpl.j[j].angle;
    }
    return true;
}

} /* namespace xpbd */
