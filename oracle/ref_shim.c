/*
 * ref_shim.c -- glue compiled together with the reference engine's own
 * sources (csrc/arc_phys.c, arc_robot.c, arc_terrain.c, arc_sim.c) by
 * oracle/Makefile into oracle/_ref/libref*.so.  It exports the same orc_*
 * functions as arc_oracle.c, with a double API whatever the build's real
 * type is, so a test can run one scenario on the restatement and on the
 * reference and compare the results bit for bit.
 *
 * Test infrastructure only; never part of the product.  The reference
 * sources are compiled where they lie and are not copied.
 */
#include <stdint.h>
#include "arc_sim.h"
#include "arc_rng.h"

#define EXPORT __attribute__((visibility("default")))
typedef double real_t; /* the build's real type (float in the FP32 build) */

/* mirrors the private struct of arc_sim.c:11-24, needed to reach its worlds */
struct arc_sim_view {
    int n;
    int threads;
    real_t dt;
    int substeps;
    int grid;
    real_t cell;
    arc_robot_info info;
    int sphere_channel[ARC_MAX_SPHERES];
    arc_world *template_world;
    arc_world *worlds;
};
#define VIEW(s) ((const struct arc_sim_view *)(s))

EXPORT int orc_real_bytes(void) { return (int)sizeof(real_t); }

EXPORT void orc_rng_stream(uint32_t seed, int n, uint32_t *out)
{
    arc_rng r;
    int i;
    arc_rng_seed(&r, seed);
    for (i = 0; i < n; ++i)
        out[i] = (uint32_t)arc_rng_u32(&r);
}
EXPORT uint32_t orc_hash32(uint32_t x) { return (uint32_t)arc_hash32(x); }
EXPORT void orc_rng_uniform(uint32_t seed, int n, f64 *out)
{
    arc_rng r;
    int i;
    arc_rng_seed(&r, seed);
    for (i = 0; i < n; ++i)
        out[i] = (f64)arc_rng_uniform(&r);
}

/* --- single worlds ---------------------------------------------------- */
EXPORT arc_world *orc_world_new(void)
{
    arc_world *w = (arc_world *)malloc(sizeof(arc_world));
    if (w)
        arc_world_init(w);
    return w;
}
EXPORT void orc_world_free(arc_world *w) { free(w); }
EXPORT void orc_world_set_gravity(arc_world *w, f64 x, f64 y, f64 z) { w->gravity = v3(x, y, z); }
EXPORT void orc_world_set_mu(arc_world *w, f64 ms, f64 md) { w->mu_static = ms; w->mu_dynamic = md; }
EXPORT void orc_terrain_flat(arc_world *w, int nx, int ny, f64 cell, f64 height)
{
    arc_terrain_flat(&w->terrain, nx, ny, cell, height);
}
EXPORT void orc_terrain_generate(arc_world *w, int n, f64 cell, f64 difficulty, uint32_t seed)
{
    arc_rng r;
    arc_rng_seed(&r, seed);
    arc_terrain_generate(&w->terrain, n, cell, difficulty, &r);
}
EXPORT void orc_terrain_set(arc_world *w, int i, f64 h) { w->terrain.h[i] = h; }
EXPORT f64 orc_terrain_height(const arc_world *w, f64 x, f64 y) { return arc_terrain_height(&w->terrain, x, y); }
EXPORT void orc_terrain_info(const arc_world *w, f64 *out)
{
    out[0] = w->terrain.nx; out[1] = w->terrain.ny; out[2] = w->terrain.cell;
    out[3] = w->terrain.ox; out[4] = w->terrain.oy; out[5] = w->terrain.outside;
}
EXPORT int orc_add_body(arc_world *w, f64 mass, const f64 *I3, const f64 *p3, const f64 *q4)
{
    return arc_add_body(w, mass, v3(I3[0], I3[1], I3[2]), v3(p3[0], p3[1], p3[2]), qt(q4[0], q4[1], q4[2], q4[3]));
}
EXPORT int orc_add_sphere(arc_world *w, int body, const f64 *p3, f64 rad)
{
    return arc_add_sphere(w, body, v3(p3[0], p3[1], p3[2]), rad);
}
EXPORT int orc_add_hinge(arc_world *w, int parent, int child, const f64 *a3, const f64 *x3)
{
    return arc_add_hinge(w, parent, child, v3(a3[0], a3[1], a3[2]), v3(x3[0], x3[1], x3[2]));
}
EXPORT void orc_joint_set(arc_world *w, int j, const f64 *p6)
{
    arc_joint *jt = &w->joints[j];
    jt->lo = p6[0]; jt->hi = p6[1]; jt->kp = p6[2]; jt->kd = p6[3]; jt->max_torque = p6[4]; jt->target = p6[5];
}
EXPORT void orc_joint_out(const arc_world *w, int j, f64 *o)
{
    o[0] = w->joints[j].angle; o[1] = w->joints[j].rate; o[2] = w->joints[j].torque;
}
EXPORT f64 orc_joint_angle(const arc_world *w, int j) { return arc_joint_angle(w, &w->joints[j]); }
EXPORT f64 orc_joint_rate(const arc_world *w, int j) { return arc_joint_rate(w, &w->joints[j]); }
EXPORT void orc_body_get(const arc_world *w, int i, f64 *o)
{
    const arc_body *b = &w->bodies[i];
    o[0] = b->x.x; o[1] = b->x.y; o[2] = b->x.z;
    o[3] = b->q.w; o[4] = b->q.x; o[5] = b->q.y; o[6] = b->q.z;
    o[7] = b->v.x; o[8] = b->v.y; o[9] = b->v.z;
    o[10] = b->w.x; o[11] = b->w.y; o[12] = b->w.z;
}
EXPORT void orc_body_set(arc_world *w, int i, const f64 *in)
{
    arc_body *b = &w->bodies[i];
    b->x = v3(in[0], in[1], in[2]);
    b->q = qt(in[3], in[4], in[5], in[6]);
    b->v = v3(in[7], in[8], in[9]);
    b->w = v3(in[10], in[11], in[12]);
}
EXPORT void orc_body_inertia(const arc_world *w, int i, f64 *o)
{
    o[0] = w->bodies[i].inertia.x; o[1] = w->bodies[i].inertia.y; o[2] = w->bodies[i].inertia.z;
}
EXPORT void orc_world_step(arc_world *w, f64 dt, int substeps, int steps)
{
    int k;
    for (k = 0; k < steps; ++k)
        arc_world_step(w, dt, substeps);
}
EXPORT void orc_world_set_pose(arc_world *w, const f64 *p3, const f64 *q4, const f64 *angles)
{
    real_t a[ARC_MAX_JOINTS];
    int i;
    for (i = 0; i < w->n_joints; ++i)
        a[i] = angles[i];
    arc_world_set_pose(w, v3(p3[0], p3[1], p3[2]), qt(q4[0], q4[1], q4[2], q4[3]), a);
}
EXPORT f64 orc_world_energy(const arc_world *w) { return arc_world_energy(w); }
EXPORT int orc_world_unstable(const arc_world *w) { return w->unstable; }
EXPORT f64 orc_world_time(const arc_world *w) { return w->time; }
EXPORT int orc_world_counts(const arc_world *w, int *o)
{
    o[0] = w->n_bodies; o[1] = w->n_joints; o[2] = w->n_spheres; o[3] = w->n_contacts;
    return w->unstable;
}
EXPORT void orc_sphere_get(const arc_world *w, int i, f64 *o)
{
    const arc_sphere *s = &w->spheres[i];
    o[0] = s->touch; o[1] = s->anchored;
    o[2] = s->anchor.x; o[3] = s->anchor.y; o[4] = s->anchor.z;
    o[5] = s->anchor_n.x; o[6] = s->anchor_n.y; o[7] = s->anchor_n.z;
}
EXPORT void orc_inertia(int kind, f64 mass, f64 a, f64 b, f64 c, f64 *o)
{
    vec3 r = kind == 0 ? arc_inertia_sphere(mass, a)
           : kind == 1 ? arc_inertia_capsule_z(mass, a, b) : arc_inertia_box(mass, v3(a, b, c));
    o[0] = r.x; o[1] = r.y; o[2] = r.z;
}

/* --- batched sim -------------------------------------------------------- */
EXPORT arc_sim *orc_sim_create(int n, f64 dt, int substeps, int grid, f64 cell)
{
    return arc_sim_create(n, dt, substeps, grid, cell);
}
EXPORT void orc_sim_destroy(arc_sim *s) { arc_sim_destroy(s); }
EXPORT int orc_sim_num_worlds(const arc_sim *s) { return arc_sim_num_worlds(s); }
EXPORT void orc_sim_robot_info(const arc_sim *s, f64 *def, f64 *lo, f64 *hi, f64 *misc3)
{
    real_t d[ARC_ROBOT_JOINTS], l[ARC_ROBOT_JOINTS], h[ARC_ROBOT_JOINTS], m[2];
    int i;
    arc_sim_robot_info(s, d, l, h, m);
    for (i = 0; i < ARC_ROBOT_JOINTS; ++i) {
        if (def) def[i] = d[i];
        if (lo) lo[i] = l[i];
        if (hi) hi[i] = h[i];
    }
    if (misc3) {
        misc3[0] = m[0];
        misc3[1] = m[1];
        misc3[2] = VIEW(s)->info.footprint;
    }
}
EXPORT arc_world *orc_sim_world(arc_sim *s, int i)
{
    return i < 0 ? VIEW(s)->template_world : (arc_world *)arc_sim_world(s, i);
}
EXPORT void orc_sim_reset(arc_sim *s, const int *ids, int n_ids, const uint32_t *seeds, const f64 *difficulty)
{
    real_t *d = (real_t *)malloc(sizeof(real_t) * (size_t)(n_ids > 0 ? n_ids : 1));
    unsigned int *sd = (unsigned int *)malloc(sizeof(unsigned int) * (size_t)(n_ids > 0 ? n_ids : 1));
    int i;
    for (i = 0; i < n_ids; ++i) {
        d[i] = difficulty[i];
        sd[i] = seeds[i];
    }
    arc_sim_set_threads(s, 1);
    arc_sim_reset(s, ids, n_ids, sd, d);
    free(d);
    free(sd);
}
EXPORT void orc_sim_step(arc_sim *s, const f64 *targets, int threads)
{
    int n = arc_sim_num_worlds(s) * ARC_ROBOT_JOINTS, i;
    arc_sim_set_threads(s, threads);
    if (sizeof(real_t) == sizeof(f64)) {
        arc_sim_step(s, (const real_t *)(const void *)targets);
        return;
    }
    {
        real_t *t = (real_t *)malloc(sizeof(real_t) * (size_t)n);
        for (i = 0; i < n; ++i)
            t[i] = targets[i];
        arc_sim_step(s, t);
        free(t);
    }
}
EXPORT void orc_sim_observe(const arc_sim *s, f64 *root, f64 *rot, f64 *joints, f64 *contacts, int *flags)
{
    int n = arc_sim_num_worlds(s), i;
    real_t *r = (real_t *)malloc(sizeof(real_t) * (size_t)n * 64);
    real_t *ro = r, *rt = r + n * 13, *jt = rt + n * 9, *ct = jt + n * 36;
    arc_sim_observe(s, ro, rt, jt, ct, flags);
    for (i = 0; i < n * 13 && root; ++i) root[i] = ro[i];
    for (i = 0; i < n * 9 && rot; ++i) rot[i] = rt[i];
    for (i = 0; i < n * 36 && joints; ++i) joints[i] = jt[i];
    for (i = 0; i < n * 6 && contacts; ++i) contacts[i] = ct[i];
    free(r);
}
EXPORT void orc_sim_heights(const arc_sim *s, const f64 *offsets, int k, f64 *out)
{
    int n = arc_sim_num_worlds(s), i;
    real_t *o = (real_t *)malloc(sizeof(real_t) * (size_t)(2 * k + n * k + 1));
    real_t *res = o + 2 * k;
    for (i = 0; i < 2 * k; ++i)
        o[i] = offsets[i];
    arc_sim_heights(s, o, k, res);
    for (i = 0; i < n * k; ++i)
        out[i] = res[i];
    free(o);
}
EXPORT void orc_sim_terrain_at(const arc_sim *s, const int *ids, const f64 *xy, int m, f64 *out)
{
    real_t *p = (real_t *)malloc(sizeof(real_t) * (size_t)(3 * m + 1));
    int i;
    for (i = 0; i < 2 * m; ++i)
        p[i] = xy[i];
    arc_sim_terrain_at(s, ids, p, m, p + 2 * m);
    for (i = 0; i < m; ++i)
        out[i] = p[2 * m + i];
    free(p);
}
EXPORT int orc_state_size(void) { return 13 * 13 + 12 * 4 + 21 * 8 + 2; }
EXPORT void orc_sim_state(const arc_sim *s, int i, f64 *out)
{
    const arc_world *w = i < 0 ? VIEW(s)->template_world : arc_sim_world(s, i);
    int k, o = 0;
    for (k = 0; k < 13; ++k) {
        orc_body_get(w, k, out + o);
        o += 13;
    }
    for (k = 0; k < 12; ++k) {
        out[o++] = w->joints[k].target; out[o++] = w->joints[k].angle;
        out[o++] = w->joints[k].rate; out[o++] = w->joints[k].torque;
    }
    for (k = 0; k < 21; ++k) {
        orc_sphere_get(w, k, out + o);
        o += 8;
    }
    out[o++] = w->unstable;
    out[o++] = w->time;
}
EXPORT void orc_sim_terrain(const arc_sim *s, int i, f64 *out)
{
    const arc_terrain *t = &arc_sim_world(s, i)->terrain;
    int k;
    for (k = 0; k < t->nx * t->ny; ++k)
        out[k] = t->h[k];
    out[t->nx * t->ny] = t->outside;
}
EXPORT int orc_sim_num_contacts(const arc_sim *s, int i) { return arc_sim_world(s, i)->n_contacts; }
