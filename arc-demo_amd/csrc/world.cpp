/*
 * world.cpp -- the single-world ABI of arc_gpu_world.h.  The world is
 * described on the host in binary32 with the builders of model.cpp (the
 * reference's arc_world API, csrc/arc_phys.c:32-143, 749-796); a step
 * uploads the dynamic state and the constants for its substep length, runs
 * the steps on one wavefront of the current GPU (k_world_step) and downloads
 * the state.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "../../include/arc_gpu_world.h"
#include "device_state.h"
#include "model.h"
#include "world_step.h"

using namespace xpbd;

namespace {

constexpr float CELL_MIN = 1e-3f, CELL_MAX = 1e3f;
constexpr int WINDOW_MAX = 2 * GRID_MAX + 2; /* columns per axis of one sphere's cell window */
constexpr int STEPS_PER_LAUNCH = 256;

/* what one step uploads, in one copy */
struct Upload {
    WorldState st;
    WorldModel m;
    WaveK wk;
};

} /* namespace */

struct arcx_world {
    HostWorld w;
    hipStream_t stream = nullptr;
    char *dmem = nullptr; /* Upload, then GRID_MAX^2 heights */
    bool heights_dirty = true;
    char msg[256] = "";
};

static int fail(const arcx_world *cw, int code, const char *what)
{
    arcx_world *w = const_cast<arcx_world *>(cw);
    snprintf(w->msg, sizeof(w->msg), "%s", what);
    return code;
}

static int fail_hip(arcx_world *w, const char *what, hipError_t e)
{
    snprintf(w->msg, sizeof(w->msg), "%s: %s", what, hipGetErrorString(e));
    return -2;
}

static V3 vec(const double *p) { return v3((float)p[0], (float)p[1], (float)p[2]); }
static Q4 quat(const double *p) { return q4((float)p[0], (float)p[1], (float)p[2], (float)p[3]); }

/* arc_terrain_height (csrc/arc_terrain.h:42-50) over the binary32 grid */
static float height_at(const HostWorld &w, float x, float y)
{
    const float fx = floorf((x - w.t.ox) / w.t.cell), fy = floorf((y - w.t.oy) / w.t.cell);
    if (!(fx >= 0.0f && fy >= 0.0f && fx < (float)w.t.nx && fy < (float)w.t.ny))
        return w.outside;
    return w.h[(size_t)fx + (size_t)fy * w.t.nx];
}

static bool cell_ok(double cell) { return cell >= CELL_MIN && cell <= CELL_MAX; }

extern "C" {

arcx_world *arcx_world_create(void)
{
    arcx_world *w = new arcx_world;
    world_init(w->w);
    hipError_t e = hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipMalloc((void **)&w->dmem, sizeof(Upload) + sizeof(float) * GRID_MAX * GRID_MAX);
    if (e != hipSuccess) {
        fprintf(stderr, "libarcsim: arcx_world_create: %s\n", hipGetErrorString(e));
        arcx_world_destroy(w);
        return nullptr;
    }
    return w;
}

void arcx_world_destroy(arcx_world *w)
{
    if (!w)
        return;
    if (w->stream)
        (void)hipStreamSynchronize(w->stream);
    if (w->dmem)
        (void)hipFree(w->dmem);
    if (w->stream)
        (void)hipStreamDestroy(w->stream);
    delete w;
}

const char *arcx_world_message(const arcx_world *w) { return w ? w->msg : "no world"; }

int arcx_world_set_gravity(arcx_world *w, double x, double y, double z)
{
    w->w.g = v3((float)x, (float)y, (float)z);
    return 0;
}

int arcx_world_set_mu(arcx_world *w, double mu_static, double mu_dynamic)
{
    w->w.mus = (float)mu_static;
    w->w.mud = (float)mu_dynamic;
    return 0;
}

int arcx_world_terrain_flat(arcx_world *w, int nx, int ny, double cell, double height)
{
    if (nx <= 0 || ny <= 0 || !cell_ok((float)cell))
        return fail(w, -1, "arcx_world_terrain_flat: grid must be positive and the cell within 1e-3 .. 1e3 m");
    terrain_flat(w->w, nx, ny, (float)cell, (float)height);
    w->heights_dirty = true;
    return 0;
}

int arcx_world_terrain_generate(arcx_world *w, int n, double cell, double difficulty, unsigned int seed)
{
    if (n <= 0 || !cell_ok((float)cell))
        return fail(w, -1, "arcx_world_terrain_generate: grid must be positive and the cell within 1e-3 .. 1e3 m");
    HostWorld &W = w->w;
    terrain_flat(W, n, n, (float)cell, 0.0f);
    W.outside = -100.0f;
    w->heights_dirty = true;
    const float d = (float)difficulty;
    if (!(d > 0.0f))
        return 0; /* flat, and nothing drawn (csrc/arc_terrain.c:60-61) */
    const int m = W.t.nx, cells = m * m;
    /* the reset's draws of a grid of m columns: both noise lattices, then one
     * jitter per column; the joint and yaw draws after them go unused */
    std::vector<float> lat(LAT_LIMIT), jit(cells);
    float draws[NJ + 1];
    reset_draws(seed, true, m, lat.data(), jit.data(), draws);
    for (int c = 0; c < cells; ++c)
        W.h[c] = 0.03f * column_step(lat.data(), m, d, jit[c], c);
    const float base = height_at(W, 0.0f, 0.0f);
    for (int c = 0; c < cells; ++c)
        W.h[c] -= base;
    return 0;
}

int arcx_world_terrain_set(arcx_world *w, int i, double height)
{
    if (i < 0 || i >= w->w.t.nx * w->w.t.ny)
        return fail(w, -1, "arcx_world_terrain_set: column index out of range");
    w->w.h[i] = (float)height;
    w->heights_dirty = true;
    return 0;
}

double arcx_world_terrain_height(const arcx_world *w, double x, double y)
{
    return (double)height_at(w->w, (float)x, (float)y);
}

int arcx_world_terrain_info(const arcx_world *w, double out6[6])
{
    const TerrainK &t = w->w.t;
    out6[0] = t.nx; out6[1] = t.ny; out6[2] = t.cell;
    out6[3] = t.ox; out6[4] = t.oy; out6[5] = w->w.outside;
    return 0;
}

int arcx_world_add_body(arcx_world *w, double mass, const double inertia3[3], const double pos3[3], const double q4_[4])
{
    int i = add_body(w->w, (float)mass, vec(inertia3), vec(pos3), quat(q4_));
    return i < 0 ? fail(w, -1, "arcx_world_add_body: 16 bodies already") : i;
}

int arcx_world_add_sphere(arcx_world *w, int body, const double pos3[3], double radius)
{
    if (body < 0 || body >= w->w.nb)
        return fail(w, -1, "arcx_world_add_sphere: no such body");
    int i = add_sphere(w->w, body, vec(pos3), (float)radius);
    return i < 0 ? fail(w, -1, "arcx_world_add_sphere: 48 spheres already") : i;
}

int arcx_world_add_hinge(arcx_world *w, int parent, int child, const double anchor3[3], const double axis3[3])
{
    if (child < 0 || child >= w->w.nb || parent < -1)
        return fail(w, -1, "arcx_world_add_hinge: no such body");
    int j = add_hinge(w->w, parent, child, vec(anchor3), vec(axis3));
    return j < 0 ? fail(w, -1, "arcx_world_add_hinge: 16 hinges already, or the child precedes the parent") : j;
}

int arcx_world_joint_set(arcx_world *w, int j, const double p6[6])
{
    if (j < 0 || j >= w->w.nj)
        return fail(w, -1, "arcx_world_joint_set: no such joint");
    HostJoint &J = w->w.j[j];
    J.lo = (float)p6[0]; J.hi = (float)p6[1]; J.kp = (float)p6[2];
    J.kd = (float)p6[3]; J.tmax = (float)p6[4]; J.target = (float)p6[5];
    return 0;
}

int arcx_world_joint_out(const arcx_world *w, int j, double out3[3])
{
    if (j < 0 || j >= w->w.nj)
        return fail(w, -1, "arcx_world_joint_out: no such joint");
    const HostJoint &J = w->w.j[j];
    out3[0] = J.angle; out3[1] = J.rate; out3[2] = J.torque;
    return 0;
}

static double joint_read(const arcx_world *w, int j, bool angle)
{
    if (j < 0 || j >= w->w.nj)
        return (double)NAN;
    const HostJoint &J = w->w.j[j];
    const JointK K = joint_constants(J, 1.0f); /* only the frames are read */
    const Body &C = w->w.b[J.child];
    if (J.parent >= 0) {
        const Body &P = w->w.b[J.parent];
        return angle ? joint_angle<true>(&P, C, K) : joint_rate<true>(&P, C, K);
    }
    return angle ? joint_angle<false>(nullptr, C, K) : joint_rate<false>(nullptr, C, K);
}

double arcx_world_joint_angle(const arcx_world *w, int j) { return joint_read(w, j, true); }
double arcx_world_joint_rate(const arcx_world *w, int j) { return joint_read(w, j, false); }

int arcx_world_body_get(const arcx_world *w, int i, double out13[13])
{
    if (i < 0 || i >= w->w.nb)
        return fail(w, -1, "arcx_world_body_get: no such body");
    const Body &b = w->w.b[i];
    const float v[13] = { b.x.x, b.x.y, b.x.z, b.q.w, b.q.x, b.q.y, b.q.z, b.v.x, b.v.y, b.v.z, b.w.x, b.w.y, b.w.z };
    for (int k = 0; k < 13; ++k)
        out13[k] = v[k];
    return 0;
}

int arcx_world_body_set(arcx_world *w, int i, const double in13[13])
{
    if (i < 0 || i >= w->w.nb)
        return fail(w, -1, "arcx_world_body_set: no such body");
    Body &b = w->w.b[i];
    b.x = vec(in13);
    b.q = quat(in13 + 3);
    b.v = vec(in13 + 7);
    b.w = vec(in13 + 10);
    return 0;
}

int arcx_world_sphere_get(const arcx_world *w, int i, double out8[8])
{
    if (i < 0 || i >= w->w.ns)
        return fail(w, -1, "arcx_world_sphere_get: no such sphere");
    const HostSphere &s = w->w.s[i];
    out8[0] = s.touch; out8[1] = s.anchored;
    out8[2] = s.anc.x; out8[3] = s.anc.y; out8[4] = s.anc.z;
    out8[5] = s.ancn.x; out8[6] = s.ancn.y; out8[7] = s.ancn.z;
    return 0;
}

int arcx_world_set_pose(arcx_world *w, const double pos3[3], const double q4_[4], const double *angles)
{
    if (w->w.nb < 1)
        return fail(w, -1, "arcx_world_set_pose: no root body");
    float a[MAX_JOINTS];
    for (int j = 0; j < w->w.nj; ++j)
        a[j] = (float)angles[j];
    set_pose(w->w, vec(pos3), quat(q4_), a);
    return 0;
}

int arcx_world_step(arcx_world *w, double dt, int substeps, int steps)
{
    HostWorld &W = w->w;
    if (!(dt > 0.0) || substeps <= 0 || steps < 0)
        return fail(w, -1, "arcx_world_step: dt must be positive, substeps at least 1, steps not negative");
    if (steps == 0)
        return 0;
    if (!cell_ok(W.t.cell))
        return fail(w, -1, "arcx_world_step: terrain cell outside 1e-3 .. 1e3 m");
    for (int k = 0; k < W.ns; ++k) {
        const float r = W.s[k].rad;
        if (!(fabsf(r) <= (float)(WINDOW_MAX - 2) * 0.5f * W.t.cell))
            return fail(w, -1, "arcx_world_step: a sphere's cell window exceeds 130 columns per axis");
    }
    Upload *up = new Upload;
    memset(up, 0, sizeof(*up));
    if (!build_wave_tables(W, W.t.cell, up->wk)) {
        delete up;
        return fail(w, -1, "arcx_world_step: the world exceeds the wavefront kernel's tables");
    }
    /* constants for this substep length, csrc/arc_phys.c:709-724 */
    WorldModel &M = up->m;
    const float dtf = (float)dt, h = dtf / (float)substeps;
    for (int b = 0; b < W.nb; ++b)
        M.body[b] = W.bk[b];
    for (int j = 0; j < W.nj; ++j)
        M.joint[j] = joint_constants(W.j[j], h);
    for (int k = 0; k < W.ns; ++k) {
        M.sphere[k].at = W.s[k].at;
        M.sphere[k].rad = W.s[k].rad;
        M.sphere[k].body = W.s[k].body;
    }
    M.terr = W.t;
    M.g = W.g;
    M.mus = W.mus;
    M.mud = W.mud;
    M.dt = dtf;
    M.h = h;
    M.hh = h * h;
    M.inv_h = 1.0f / h;
    M.two_h = 2.0f / h;
    M.substeps = substeps;
    M.outside = W.outside;
    WorldState &S = up->st;
    for (int b = 0; b < W.nb; ++b)
        S.b[b] = W.b[b];
    for (int j = 0; j < W.nj; ++j) {
        S.target[j] = W.j[j].target;
        S.angle[j] = W.j[j].angle;
        S.rate[j] = W.j[j].rate;
        S.torque[j] = W.j[j].torque;
    }
    for (int k = 0; k < W.ns; ++k) {
        S.touch[k] = W.s[k].touch;
        S.anchored[k] = W.s[k].anchored;
        S.anc[k] = W.s[k].anc;
        S.ancn[k] = W.s[k].ancn;
    }
    S.unstable = W.unstable;
    S.time = W.time;

    Upload *d_up = (Upload *)w->dmem;
    float *d_h = (float *)(w->dmem + sizeof(Upload));
    hipError_t e = hipMemcpyAsync(d_up, up, sizeof(Upload), hipMemcpyHostToDevice, w->stream);
    if (e == hipSuccess && w->heights_dirty)
        e = hipMemcpyAsync(d_h, W.h.data(), sizeof(float) * W.h.size(), hipMemcpyHostToDevice, w->stream);
    for (int done = 0; e == hipSuccess && done < steps; done += STEPS_PER_LAUNCH) {
        const int k = steps - done < STEPS_PER_LAUNCH ? steps - done : STEPS_PER_LAUNCH;
        e = launch_world_step(&d_up->st, &d_up->m, &d_up->wk, d_h, k, w->stream);
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(&up->st, &d_up->st, sizeof(WorldState), hipMemcpyDeviceToHost, w->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(w->stream);
    if (e != hipSuccess) {
        delete up;
        return fail_hip(w, "arcx_world_step", e);
    }
    w->heights_dirty = false;
    for (int b = 0; b < W.nb; ++b) {
        W.b[b].x = S.b[b].x;
        W.b[b].q = S.b[b].q;
        W.b[b].v = S.b[b].v;
        W.b[b].w = S.b[b].w;
        bsync<KGEN>(W.b[b], W.bk[b]);
    }
    for (int j = 0; j < W.nj; ++j) {
        W.j[j].angle = S.angle[j];
        W.j[j].rate = S.rate[j];
        W.j[j].torque = S.torque[j];
    }
    for (int k = 0; k < W.ns; ++k) {
        W.s[k].touch = S.touch[k];
        W.s[k].anchored = S.anchored[k];
        W.s[k].anc = S.anc[k];
        W.s[k].ancn = S.ancn[k];
    }
    W.unstable = S.unstable;
    W.time = S.time;
    const int overflow = S.overflow;
    delete up;
    if (overflow)
        return fail(w, -2, "arcx_world_step: a contact list overflowed its slots");
    return 0;
}

double arcx_world_energy(const arcx_world *w)
{
    /* csrc/arc_phys.c:784-796 in binary32, term by term in the reference's order */
    const HostWorld &W = w->w;
    float e = 0.0f;
    for (int i = 0; i < W.nb; ++i) {
        const Body &b = W.b[i];
        const BodyK &k = W.bk[i];
        const V3 wl = qirot(b.q, b.w);
        e += 0.5f * k.mass * len2(b.v);
        e += 0.5f * dot(wl, had(wl, k.I));
        e -= k.mass * dot(W.g, b.x);
    }
    return (double)e;
}

int arcx_world_unstable(const arcx_world *w) { return w->w.unstable; }
double arcx_world_time(const arcx_world *w) { return (double)w->w.time; }

int arcx_world_counts(const arcx_world *w, int out3[3])
{
    out3[0] = w->w.nb;
    out3[1] = w->w.nj;
    out3[2] = w->w.ns;
    return 0;
}

} /* extern "C" */
