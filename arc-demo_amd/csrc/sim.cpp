/*
 * sim.cpp -- host side of libarcsim: the arc_sim.h C ABI over the device
 * engine (kernels.hip).  Builds the robot once in binary32 on the host,
 * uploads its constants, keeps every world's state resident in HBM and moves
 * only the ABI's arrays across PCIe.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <vector>

#include "model.h"
#include "reset_engine.h"
#include "noalloc.h"
#include "sim_internal.h"
#include "visuals.h"

void set_error(const arc_sim *cs, int code, const char *what, hipError_t e)
{
    arc_sim *s = const_cast<arc_sim *>(cs);
    if (!s->err) {
        s->err = code;
        snprintf(s->msg, sizeof(s->msg), "%s: %s", what, e == hipSuccess ? "invalid argument" : hipGetErrorString(e));
    }
    fprintf(stderr, "libarcsim: %s: %s\n", what, e == hipSuccess ? "invalid argument" : hipGetErrorString(e));
}


static bool ensure_io(arc_sim *s, size_t doubles)
{
    if (doubles <= s->d_io_doubles)
        return true;
    if (s->d_io) {
        (void)hipStreamSynchronize(s->stream); /* a queued step may still read it */
        (void)hipFree(s->d_io);
    }
    s->d_io = nullptr;
    s->d_io_doubles = 0;
    hipError_t e = hipMalloc((void **)&s->d_io, doubles * sizeof(double));
    if (e != hipSuccess) {
        set_error(s, (int)e, "device staging", e);
        return false;
    }
    s->d_io_doubles = doubles;
    return true;
}
static bool ensure_aux(arc_sim *s, size_t bytes)
{
    if (bytes <= s->d_aux_bytes)
        return true;
    if (s->d_aux)
        (void)hipFree(s->d_aux);
    s->d_aux = nullptr;
    s->d_aux_bytes = 0;
    hipError_t e = hipMalloc((void **)&s->d_aux, bytes);
    if (e != hipSuccess) {
        set_error(s, (int)e, "device staging", e);
        return false;
    }
    s->d_aux_bytes = bytes;
    return true;
}
static bool ensure_pin(arc_sim *s, size_t bytes)
{
    if (bytes <= s->h_pin_bytes)
        return true;
    if (s->h_pin) {
        (void)hipStreamSynchronize(s->stream); /* a queued upload may still read it */
        (void)hipHostFree(s->h_pin);
    }
    s->h_pin = nullptr;
    s->h_pin_bytes = 0;
    hipError_t e = hipHostMalloc((void **)&s->h_pin, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        set_error(s, (int)e, "pinned staging", e);
        return false;
    }
    s->h_pin_bytes = bytes;
    return true;
}

/* [p, p + bytes) lies in one allocation of this sim's GPU: the step kernels
 * read targets there, and a host address, or a range that runs past the end
 * of its allocation, would fault the launch */
bool device_memory(const arc_sim *s, const void *p, size_t bytes)
{
    hipPointerAttribute_t a{};
    if (!p || hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError(); /* clear the lookup's error */
        return false;
    }
    if (!(a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged) || a.device != s->device)
        return false;
    if (bytes == 0)
        return true;
    hipDeviceptr_t base = nullptr;
    size_t size = 0;
    if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    const uintptr_t lo = (uintptr_t)base, at = (uintptr_t)p;
    return at >= lo && bytes <= size - (at - lo);
}

/* the configured form of the step kernel over all worlds */
hipError_t launch_form(arc_sim *s, const float *tf)
{
    note_contacts(s);
    if (s->form == 1)
        return launch_step_wave(s->dev, s->d_model, s->d_wave, nullptr, tf, 0, s->stream);
    return launch_lane_all(s, tf);
}

extern "C" {

int arcx_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return 0;
    return n;
}

arc_sim *arc_sim_create(int n_worlds, double dt, int substeps, int grid, double cell)
{
    if (n_worlds < 0 || substeps <= 0 || grid <= 0 || !(cell > 0.0)) {
        fprintf(stderr, "libarcsim: arc_sim_create: invalid arguments\n");
        return nullptr;
    }
    if (arcx_device_count() <= 0) {
        fprintf(stderr, "libarcsim: arc_sim_create: no HIP device available; this library runs only on the GPU\n");
        return nullptr;
    }
    arc_sim *s = new (std::nothrow) arc_sim();
    if (!s)
        return nullptr;
    s->n = n_worlds;
    s->substeps = substeps;
    s->dt_d = dt;
    s->cell_d = cell;
    s->block = 64;
    s->grid_blocks = 0;
    s->tile = 16;
    s->persistent = 0;
    (void)hipGetDevice(&s->device);
    /* the device address of the allocation-free entry point's static scratch,
     * looked up once per GPU with that GPU current; a failure here disables
     * only that entry point (it then reports an error) */
    {
        const hipError_t se = noalloc_scratch_init(s->device);
        if (se != hipSuccess)
            fprintf(stderr, "libarcsim: arc_sim_create: static scratch lookup failed (%s)\n", hipGetErrorString(se));
    }

    if (!sim_setup(s->su, dt, substeps, grid, cell)) {
        delete s;
        return nullptr;
    }
    {
        /* the spheres' column windows divide by the column size: a
         * three-instruction quotient where the device check proves it exact
         * and the numerators stay in its domain (radii and a non-zero grid
         * origin in [2^-60, 2^90), see sphere_window) */
        Model &Mq = s->su.model;
        auto in_range = [](float v) { return fabsf(v) >= 0x1p-60f && fabsf(v) < 0x1p90f; };
        bool ok = (Mq.terr.ox == 0.0f || in_range(Mq.terr.ox)) && (Mq.terr.oy == 0.0f || in_range(Mq.terr.oy));
        for (int k = 0; k < NS; ++k)
            ok = ok && in_range(Mq.sphere[k].rad);
        Mq.terr.inv_cell = 1.0f / Mq.terr.cell;
        Mq.terr.fast_cell = ok && quotient_checked(Mq.terr.cell);
    }
    const Model &M = s->su.model;
    if (!build_wave_tables(s->su.tmpl, (float)cell, s->wave)) {
        fprintf(stderr, "libarcsim: world exceeds the wavefront kernel's tables\n");
        delete s;
        return nullptr;
    }
    hipError_t e;
    const size_t jtab_bytes = s->su.jtab.size() * sizeof(uint32_t);
    ResetK rk_dev = s->su.rk; /* the device copy points at the device tables */
    if ((e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) == hipSuccess
        && (e = hipEventCreateWithFlags(&s->up_done, hipEventDisableTiming)) == hipSuccess
        && (e = dev_alloc(s->dev, n_worlds, s->su.grid, M.cslots)) == hipSuccess
        && (e = hipMalloc((void **)&s->d_model, sizeof(Model))) == hipSuccess
        && (e = hipMalloc((void **)&s->d_reset, sizeof(ResetK))) == hipSuccess
        && (e = hipMalloc((void **)&s->d_jtab, jtab_bytes)) == hipSuccess
        && (e = hipMalloc((void **)&s->d_tmpl, sizeof(TemplateK))) == hipSuccess
        && (e = hipMalloc((void **)&s->d_dump, 1024 * sizeof(float))) == hipSuccess
        && (e = hipMalloc((void **)&s->d_wave, sizeof(WaveK))) == hipSuccess) {
        rk_dev.jtab = s->d_jtab;
        if ((e = hipMemcpyAsync(s->d_jtab, s->su.jtab.data(), jtab_bytes, hipMemcpyHostToDevice, s->stream))
                == hipSuccess
            && (e = hipMemcpyAsync(s->d_wave, &s->wave, sizeof(WaveK), hipMemcpyHostToDevice, s->stream)) == hipSuccess
            && (e = hipMemcpyAsync(s->d_model, &M, sizeof(Model), hipMemcpyHostToDevice, s->stream)) == hipSuccess
            && (e = hipMemcpyAsync(s->d_reset, &rk_dev, sizeof(ResetK), hipMemcpyHostToDevice, s->stream)) == hipSuccess
            && (e = hipMemcpyAsync(s->d_tmpl, &s->su.tk, sizeof(TemplateK), hipMemcpyHostToDevice, s->stream))
                   == hipSuccess)
            e = launch_init_template(s->dev, s->d_tmpl, s->su.tmpl.outside, s->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(s->stream);
    }
    if (e != hipSuccess) {
        fprintf(stderr, "libarcsim: arc_sim_create: %s\n", hipGetErrorString(e));
        arc_sim_destroy(s);
        return nullptr;
    }
    /* task defaults of arc/task.py:57-59 */
    s->task.half_extent = 0.5 * s->su.grid * cell;
    s->task.void_z = -2.0;
    s->task.flip_up_z = 0.3;
    s->task.beached_steps = 30;
    s->task.n = n_worlds;
    s->reset_blocks = 1024;
    {
        size_t N = (size_t)(n_worlds > 0 ? n_worlds : 1);
        size_t bytes = N * (sizeof(int) * 3 + sizeof(uint32_t) * 2 + sizeof(float)) + 64 + 6 * sizeof(unsigned long long);
        char *p = nullptr;
        if ((e = hipMalloc((void **)&p, bytes)) != hipSuccess || (e = hipMemsetAsync(p, 0, bytes, s->stream)) != hipSuccess
            || (e = hipStreamSynchronize(s->stream)) != hipSuccess) {
            fprintf(stderr, "libarcsim: arc_sim_create: %s\n", hipGetErrorString(e));
            arc_sim_destroy(s);
            return nullptr;
        }
        s->ts_mem = p;
        unsigned long long *st = (unsigned long long *)p;
        s->ts.stats = st;
        s->ts.count = (int *)(st + 6);
        s->ts.torso_steps = (int *)(p + 64 + 6 * sizeof(unsigned long long));
        s->ts.cause = s->ts.torso_steps + N;
        s->ts.list = s->ts.cause + N;
        s->ts.episode = (uint32_t *)(s->ts.list + N);
        s->ts.seeds = s->ts.episode + N;
        s->ts.diff = (float *)(s->ts.seeds + N);
    }
    /* default form: four lanes per world, the fastest untuned form at every
     * world count measured (64 to 65,536); its contact lists are made here so
     * that a failed allocation is a failed create, as in the reference.  A
     * robot whose contact windows exceed the group forms' cap keeps the
     * one-world-per-lane form. */
    e = ensure_leg(s);
    if (e == hipSuccess) {
        s->form = 2;
    } else if (e != hipErrorNotSupported) {
        fprintf(stderr, "libarcsim: arc_sim_create: %s\n", hipGetErrorString(e));
        arc_sim_destroy(s);
        return nullptr;
    }
    return s;
}

void arc_sim_destroy(arc_sim *s)
{
    if (!s)
        return;
    if (s->stream)
        (void)hipStreamSynchronize(s->stream);
    batch_free(s);
    if (s->dev.lcbuf)
        (void)hipFree(s->dev.lcbuf);
    dev_free(s->dev);
    if (s->d_model) (void)hipFree(s->d_model);
    if (s->d_reset) (void)hipFree(s->d_reset);
    if (s->d_jtab) (void)hipFree(s->d_jtab);
    if (s->d_tmpl) (void)hipFree(s->d_tmpl);
    if (s->d_dump) (void)hipFree(s->d_dump);
    if (s->d_wave) (void)hipFree(s->d_wave);
    if (s->ts_mem) (void)hipFree(s->ts_mem);
    if (s->d_io) (void)hipFree(s->d_io);
    if (s->d_aux) (void)hipFree(s->d_aux);
    if (s->h_pin) (void)hipHostFree(s->h_pin);
    if (s->up_done) (void)hipEventDestroy(s->up_done);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    free(s->mirror);
    delete s;
}

void arc_sim_set_threads(arc_sim *s, int threads)
{
    (void)s;
    (void)threads;
}

int arc_sim_num_worlds(const arc_sim *s) { return s->n; }
int arc_sim_num_joints(void) { return ROBOT_JOINTS; }
int arc_sim_num_legs(void) { return ROBOT_LEGS; }

void arc_sim_robot_info(const arc_sim *s, double *def, double *lo, double *hi, double *misc)
{
    for (int i = 0; i < ROBOT_JOINTS; ++i) {
        if (def) def[i] = s->su.info.def[i];
        if (lo) lo[i] = s->su.info.lo[i];
        if (hi) hi[i] = s->su.info.hi[i];
    }
    if (misc) {
        misc[0] = s->su.info.stand;
        misc[1] = s->su.info.mass;
    }
}

void arc_sim_reset(arc_sim *s, const int *ids, int n_ids, const unsigned int *seeds, const double *difficulty)
{
    if (n_ids <= 0)
        return;
    /* last entry per world wins (the reference resets them in list order) */
    std::vector<int> last(s->n, -1);
    int bad = 0, steep = 0;
    for (int k = 0; k < n_ids; ++k) {
        if (ids[k] < 0 || ids[k] >= s->n) {
            ++bad;
            continue;
        }
        /* the device terrain holds each column as an int8 count of 0.03 m
         * steps, |F| <= 21 d + 1 before the re-basing: difficulties up to
         * MAX_DIFFICULTY fit (NaN and d <= 0 give the flat terrain) */
        if (difficulty[k] > MAX_DIFFICULTY) {
            ++steep;
            continue;
        }
        last[ids[k]] = k;
    }
    if (bad)
        set_error(s, -1, "arc_sim_reset: world id out of range (skipped)", hipSuccess);
    if (steep)
        set_error(s, -1, "arc_sim_reset: difficulty above 6, beyond the device terrain's range (skipped)", hipSuccess);
    std::vector<int> kid;
    kid.reserve(n_ids);
    for (int k = 0; k < n_ids; ++k)
        if (ids[k] >= 0 && ids[k] < s->n && last[ids[k]] == k)
            kid.push_back(k);
    int m = (int)kid.size();
    if (!m)
        return;
    size_t bytes = (size_t)m * (sizeof(int) + sizeof(uint32_t) + sizeof(double));
    if (!ensure_aux(s, bytes) || !ensure_pin(s, bytes))
        return;
    int *hid = (int *)s->h_pin;
    uint32_t *hseed = (uint32_t *)(hid + m);
    double *hdiff = (double *)(s->h_pin + (size_t)m * (sizeof(int) + sizeof(uint32_t)));
    hipStreamSynchronize(s->stream);
    for (int t = 0; t < m; ++t) {
        hid[t] = ids[kid[t]];
        hseed[t] = seeds[kid[t]];
        hdiff[t] = difficulty[kid[t]];
    }
    int *did = (int *)s->d_aux;
    uint32_t *dseed = (uint32_t *)(did + m);
    double *ddiff = (double *)(s->d_aux + (size_t)m * (sizeof(int) + sizeof(uint32_t)));
    /* the double block must be 8-byte aligned: m*8 bytes precede it */
    HIPCHK(s, hipMemcpyAsync(s->d_aux, s->h_pin, bytes, hipMemcpyHostToDevice, s->stream), "arc_sim_reset upload");
    HIPCHK(s, launch_reset(s->dev, s->d_model, s->d_reset, did, dseed, ddiff, nullptr, m, nullptr, 4096, s->stream),
           "arc_sim_reset launch");
    HIPCHK(s, hipStreamSynchronize(s->stream), "arc_sim_reset");
}

void arc_sim_step(arc_sim *s, const double *targets)
{
    size_t nt = (size_t)s->n * NJ;
    if (!nt)
        return;
    if (!ensure_io(s, nt) || !ensure_pin(s, nt * sizeof(float)))
        return;
    /* the staging buffer is free again once the previous step's upload is done
     * (that step itself may still be running) */
    if (s->up_pending) {
        HIPCHK(s, hipEventSynchronize(s->up_done), "arc_sim_step");
        s->up_pending = false;
    }
    /* the engine reads binary32 targets; rounding before the clamp to binary32
     * limits gives the clamp of the rounded value, as the FP32 engine computes */
    float *hf = (float *)s->h_pin;
    for (size_t k = 0; k < nt; ++k)
        hf[k] = (float)targets[k];
    HIPCHK(s, hipMemcpyAsync(s->d_io, s->h_pin, nt * sizeof(float), hipMemcpyHostToDevice, s->stream), "arc_sim_step upload");
    HIPCHK(s, hipEventRecord(s->up_done, s->stream), "arc_sim_step upload");
    s->up_pending = true;
    HIPCHK(s, launch_form(s, (const float *)s->d_io), "arc_sim_step launch");
    /* returns with the step queued on the sim's stream: every entry point that
     * reads results on the host, or rewrites what the step reads, waits for
     * that stream first, so the next call's conversion overlaps this step */
}

void arc_sim_observe(const arc_sim *cs, double *root, double *rot, double *joints, double *contacts, int *flags)
{
    arc_sim *s = const_cast<arc_sim *>(cs);
    size_t n = (size_t)s->n;
    if (!n)
        return;
    const size_t per = 13 + 9 + 36 + 6 + 1; /* flags as one double slot */
    if (!ensure_io(s, n * per) || !ensure_pin(s, n * per * sizeof(double)))
        return;
    double *d_root = s->d_io, *d_rot = d_root + n * 13, *d_j = d_rot + n * 9, *d_c = d_j + n * 36;
    int *d_f = (int *)(d_c + n * 6);
    HIPCHK(s, launch_observe(s->dev, s->d_model, root ? d_root : nullptr, rot ? d_rot : nullptr,
                             joints ? d_j : nullptr, contacts ? d_c : nullptr, flags ? d_f : nullptr, s->stream),
           "arc_sim_observe launch");
    /* only the requested outputs cross PCIe (each is one contiguous section) */
    double *h = (double *)s->h_pin;
    const struct { const void *want; size_t off, bytes; } part[5] = {
        { root, 0, n * 13 * sizeof(double) },          { rot, n * 13, n * 9 * sizeof(double) },
        { joints, n * 22, n * 36 * sizeof(double) },   { contacts, n * 58, n * 6 * sizeof(double) },
        { flags, n * 64, n * sizeof(int) } };
    for (const auto &p : part)
        if (p.want)
            HIPCHK(s, hipMemcpyAsync(h + p.off, s->d_io + p.off, p.bytes, hipMemcpyDeviceToHost, s->stream),
                   "arc_sim_observe download");
    HIPCHK(s, hipStreamSynchronize(s->stream), "arc_sim_observe");
    if (root) memcpy(root, h, n * 13 * sizeof(double));
    if (rot) memcpy(rot, h + n * 13, n * 9 * sizeof(double));
    if (joints) memcpy(joints, h + n * 22, n * 36 * sizeof(double));
    if (contacts) memcpy(contacts, h + n * 58, n * 6 * sizeof(double));
    if (flags) memcpy(flags, h + n * 64, n * sizeof(int));
}

void arc_sim_heights(const arc_sim *cs, const double *offsets, int k, double *out)
{
    arc_sim *s = const_cast<arc_sim *>(cs);
    size_t n = (size_t)s->n;
    if (!n || k <= 0)
        return;
    size_t outd = n * (size_t)k;
    if (!ensure_io(s, outd) || !ensure_aux(s, (size_t)k * 2 * sizeof(double))
        || !ensure_pin(s, std::max(outd, (size_t)k * 2) * sizeof(double)))
        return;
    HIPCHK(s, hipStreamSynchronize(s->stream), "arc_sim_heights");
    memcpy(s->h_pin, offsets, (size_t)k * 2 * sizeof(double));
    HIPCHK(s, hipMemcpyAsync(s->d_aux, s->h_pin, (size_t)k * 2 * sizeof(double), hipMemcpyHostToDevice, s->stream),
           "arc_sim_heights upload");
    HIPCHK(s, launch_heights(s->dev, s->d_model, (const double *)s->d_aux, k, s->d_io, s->stream), "arc_sim_heights launch");
    HIPCHK(s, hipMemcpyAsync(s->h_pin, s->d_io, outd * sizeof(double), hipMemcpyDeviceToHost, s->stream),
           "arc_sim_heights download");
    HIPCHK(s, hipStreamSynchronize(s->stream), "arc_sim_heights");
    memcpy(out, s->h_pin, outd * sizeof(double));
}

void arc_sim_terrain_at(const arc_sim *cs, const int *ids, const double *xy, int m, double *out)
{
    arc_sim *s = const_cast<arc_sim *>(cs);
    if (m <= 0)
        return;
    std::vector<int> idx(m);
    int bad = 0;
    for (int i = 0; i < m; ++i) {
        idx[i] = ids[i];
        if (ids[i] < 0 || ids[i] >= s->n) {
            idx[i] = 0;
            ++bad;
        }
    }
    if (bad || s->n == 0)
        set_error(s, -1, "arc_sim_terrain_at: world id out of range", hipSuccess);
    if (s->n == 0)
        return;
    size_t bytes = (size_t)m * (2 * sizeof(double) + sizeof(int) + sizeof(double));
    if (!ensure_aux(s, bytes) || !ensure_pin(s, bytes) || !ensure_io(s, (size_t)m))
        return;
    HIPCHK(s, hipStreamSynchronize(s->stream), "arc_sim_terrain_at");
    memcpy(s->h_pin, xy, (size_t)m * 2 * sizeof(double));
    memcpy(s->h_pin + (size_t)m * 2 * sizeof(double), idx.data(), (size_t)m * sizeof(int));
    HIPCHK(s, hipMemcpyAsync(s->d_aux, s->h_pin, (size_t)m * (2 * sizeof(double) + sizeof(int)), hipMemcpyHostToDevice,
                             s->stream), "arc_sim_terrain_at upload");
    HIPCHK(s, launch_terrain_at(s->dev, s->d_model, (const int *)(s->d_aux + (size_t)m * 2 * sizeof(double)),
                                (const double *)s->d_aux, m, s->d_io, s->stream), "arc_sim_terrain_at launch");
    HIPCHK(s, hipMemcpyAsync(s->h_pin, s->d_io, (size_t)m * sizeof(double), hipMemcpyDeviceToHost, s->stream),
           "arc_sim_terrain_at download");
    HIPCHK(s, hipStreamSynchronize(s->stream), "arc_sim_terrain_at");
    memcpy(out, s->h_pin, (size_t)m * sizeof(double));
    for (int i = 0; i < m; ++i)
        if (ids[i] < 0 || ids[i] >= s->n)
            out[i] = NAN;
}

/* -------------------------------------------------------------------- */

void arcx_rng_stream(unsigned int seed, long long n, unsigned int *out)
{
    Rng r;
    r.seed(seed);
    for (long long i = 0; i < n; ++i)
        out[i] = r.next();
}

void *arcx_device_alloc(size_t bytes)
{
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess)
        return nullptr;
    return p;
}

int arcx_device_free(void *p) { return p ? (int)hipFree(p) : 0; }

int arcx_memcpy(void *dst, const void *src, size_t bytes, int to_device)
{
    hipMemcpyKind k = to_device == 1 ? hipMemcpyHostToDevice : to_device == 2 ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    hipError_t e = hipMemcpy(dst, src, bytes, k);
    return (int)e;
}

int arcx_state_size(void) { return NB * 13 + NJ * 4 + NS * 8 + 2; }

int arcx_state(const arc_sim *cs, int i, double *out)
{
    arc_sim *s = const_cast<arc_sim *>(cs);
    const int ns = arcx_state_size();
    if (i < 0) { /* the template, from the host builder */
        int o = 0;
        for (int b = 0; b < NB; ++b) {
            const Body &B = s->su.tmpl.b[b];
            double v[13] = { B.x.x, B.x.y, B.x.z, B.q.w, B.q.x, B.q.y, B.q.z, B.v.x, B.v.y, B.v.z, B.w.x, B.w.y, B.w.z };
            for (int c = 0; c < 13; ++c) out[o++] = v[c];
        }
        for (int j = 0; j < NJ; ++j) {
            const HostJoint &J = s->su.tmpl.j[j];
            out[o++] = J.target; out[o++] = J.angle; out[o++] = J.rate; out[o++] = J.torque;
        }
        for (int k = 0; k < NS; ++k) {
            const HostSphere &S = s->su.tmpl.s[k];
            double v[8] = { (double)S.touch, (double)S.anchored, S.anc.x, S.anc.y, S.anc.z, S.ancn.x, S.ancn.y, S.ancn.z };
            for (int c = 0; c < 8; ++c) out[o++] = v[c];
        }
        out[o++] = s->su.tmpl.unstable;
        out[o++] = s->su.tmpl.time;
        return 0;
    }
    if (i >= s->n)
        return -1;
    float buf[1024];
    HIPCHK_RET(s, launch_dump(s->dev, i, s->d_dump, s->stream), "arcx_state", -1);
    HIPCHK_RET(s, hipMemcpyAsync(buf, s->d_dump, ns * sizeof(float), hipMemcpyDeviceToHost, s->stream), "arcx_state", -1);
    HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arcx_state", -1);
    for (int k = 0; k < ns; ++k)
        out[k] = buf[k];
    return 0;
}

int arcx_terrain(const arc_sim *cs, int i, double *out)
{
    arc_sim *s = const_cast<arc_sim *>(cs);
    if (i < 0 || i >= s->n)
        return -1;
    const DevWorlds &d = s->dev;
    std::vector<int8_t> q(d.tcells);
    float bo[2];
    HIPCHK_RET(s, hipMemcpyAsync(q.data(), d.tq + (size_t)i * d.tcells, d.tcells, hipMemcpyDeviceToHost, s->stream),
               "arcx_terrain", -1);
    HIPCHK_RET(s, hipMemcpyAsync(&bo[0], d.tbase + i, sizeof(float), hipMemcpyDeviceToHost, s->stream), "arcx_terrain", -1);
    HIPCHK_RET(s, hipMemcpyAsync(&bo[1], d.outside + i, sizeof(float), hipMemcpyDeviceToHost, s->stream), "arcx_terrain", -1);
    HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arcx_terrain", -1);
    TerrainRef r;
    r.q = q.data() + TPAD + TPAD * d.tpitch;
    r.pitch = d.tpitch;
    r.base = bo[0];
    r.outside = bo[1];
    const int n = s->su.grid;
    for (int iy = 0; iy < n; ++iy)
        for (int ix = 0; ix < n; ++ix)
            out[ix + iy * n] = terrain_cell(s->su.model.terr, r, ix, iy);
    out[n * n] = bo[1];
    return 0;
}

int arcx_step_device(arc_sim *s, const float *targets_dev)
{
    if (!device_memory(s, targets_dev, (size_t)s->n * NJ * sizeof(float))) {
        set_error(s, -1, "arcx_step_device: targets are not device memory of the sim's GPU", hipSuccess);
        return -1;
    }
    HIPCHK_RET(s, launch_form(s, targets_dev), "arcx_step_device", -1);
    return 0;
}

/* steps back to back with device targets[k * step_floats ...]; HIP events
 * around each launch on the engine's stream give the kernel time */
int arcx_time_steps(arc_sim *s, const float *targets_dev, int steps, long long step_floats, double *kernel_ms,
                    double *wall_ms)
{
    if (steps < 0 || step_floats < 0
        || !device_memory(s, targets_dev,
                          ((size_t)(steps > 0 ? steps - 1 : 0) * (size_t)step_floats + (size_t)s->n * NJ)
                              * sizeof(float))) {
        set_error(s, -1, "arcx_time_steps: targets are not device memory of the sim's GPU", hipSuccess);
        return -1;
    }
    hipEvent_t e0, e1, t0, t1;
    HIPCHK_RET(s, hipEventCreate(&e0), "arcx_time_steps", -1);
    HIPCHK_RET(s, hipEventCreate(&e1), "arcx_time_steps", -1);
    HIPCHK_RET(s, hipEventCreate(&t0), "arcx_time_steps", -1);
    HIPCHK_RET(s, hipEventCreate(&t1), "arcx_time_steps", -1);
    double total = 0.0;
    HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arcx_time_steps", -1);
    HIPCHK_RET(s, hipEventRecord(t0, s->stream), "arcx_time_steps", -1);
    for (int k = 0; k < steps; ++k) {
        HIPCHK_RET(s, hipEventRecord(e0, s->stream), "arcx_time_steps", -1);
        HIPCHK_RET(s, launch_form(s, targets_dev + (size_t)k * step_floats), "arcx_time_steps", -1);
        HIPCHK_RET(s, hipEventRecord(e1, s->stream), "arcx_time_steps", -1);
        HIPCHK_RET(s, hipEventSynchronize(e1), "arcx_time_steps", -1);
        float ms = 0.0f;
        HIPCHK_RET(s, hipEventElapsedTime(&ms, e0, e1), "arcx_time_steps", -1);
        total += ms;
    }
    HIPCHK_RET(s, hipEventRecord(t1, s->stream), "arcx_time_steps", -1);
    HIPCHK_RET(s, hipEventSynchronize(t1), "arcx_time_steps", -1);
    float wall = 0.0f;
    HIPCHK_RET(s, hipEventElapsedTime(&wall, t0, t1), "arcx_time_steps", -1);
    if (kernel_ms)
        *kernel_ms = total;
    if (wall_ms)
        *wall_ms = wall;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    return 0;
}

int arcx_task_config(arc_sim *s, int beached_steps, double flip_up_z, double void_z)
{
    s->task.beached_steps = beached_steps;
    s->task.flip_up_z = flip_up_z;
    s->task.void_z = void_z;
    return 0;
}

/* step, fall rules, auto-reset of the fallen worlds: no host round trip */
int arcx_step_task(arc_sim *s, const float *targets_dev)
{
    if (!device_memory(s, targets_dev, (size_t)s->n * NJ * sizeof(float))) {
        set_error(s, -1, "arcx_step_task: targets are not device memory of the sim's GPU", hipSuccess);
        return -1;
    }
    HIPCHK_RET(s, launch_fill_int(s->ts.count, 0, 1, s->stream), "arcx_step_task", -1);
    HIPCHK_RET(s, launch_form(s, targets_dev), "arcx_step_task step", -1);
    HIPCHK_RET(s, launch_falls(s->dev, s->d_model, s->task, s->ts, nullptr, s->stream), "arcx_step_task falls", -1);
    HIPCHK_RET(s, launch_reset(s->dev, s->d_model, s->d_reset, s->ts.list, s->ts.seeds, nullptr, s->ts.diff, 0,
                               s->ts.count, s->reset_blocks, s->stream), "arcx_step_task reset", -1);
    return 0;
}

int arcx_task_causes(const arc_sim *cs, int *out)
{
    arc_sim *s = const_cast<arc_sim *>(cs);
    HIPCHK_RET(s, hipMemcpyAsync(out, s->ts.cause, (size_t)s->n * sizeof(int), hipMemcpyDeviceToHost, s->stream),
               "arcx_task_causes", -1);
    HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arcx_task_causes", -1);
    return 0;
}

int arcx_task_stats(const arc_sim *cs, long long *out5)
{
    arc_sim *s = const_cast<arc_sim *>(cs);
    unsigned long long v[5];
    HIPCHK_RET(s, hipMemcpyAsync(v, s->ts.stats, sizeof(v), hipMemcpyDeviceToHost, s->stream), "arcx_task_stats", -1);
    HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arcx_task_stats", -1);
    for (int k = 0; k < 5; ++k)
        out5[k] = (long long)v[k];
    return 0;
}

int arcx_task_episodes(const arc_sim *cs, unsigned int *out)
{
    arc_sim *s = const_cast<arc_sim *>(cs);
    HIPCHK_RET(s, hipMemcpyAsync(out, s->ts.episode, (size_t)s->n * sizeof(uint32_t), hipMemcpyDeviceToHost, s->stream),
               "arcx_task_episodes", -1);
    HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arcx_task_episodes", -1);
    return 0;
}

int arcx_sync(arc_sim *s)
{
    HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arcx_sync", -1);
    return s->err;
}

int arcx_set_form(arc_sim *s, int form)
{
    if (form < 0 || form > 4)
        return -1;
    if (form >= 2)
        HIPCHK_RET(s, ensure_leg(s), "arcx_set_form: group-form buffers", -1);
    s->form = form;
    if (form == 4) { /* blocks of 128 or 256 threads, no work counter */
        if (s->block == 64)
            s->block = 256;
        s->persistent = 0;
    }
    return 0;
}

int arcx_set_launch(arc_sim *s, int block, int grid_blocks)
{
    if ((block != 64 && block != 128 && block != 256) || grid_blocks < 0 || (s->form == 4 && block == 64))
        return -1;
    s->block = block;
    s->grid_blocks = grid_blocks;
    return 0;
}

double arcx_bytes_per_world(const arc_sim *s)
{
    /* the world arrays plus the group forms' contact lists (held once one is used) */
    const Model &M = s->su.model;
    const double leg = (double)std::max(leg_world_slots(M, 1), leg_world_slots(M, 2)) * 6 * sizeof(float);
    return (double)dev_bytes_per_world(s->su.grid, M.cslots) + leg;
}

int arcx_world_error(const arc_sim *cs, int i)
{
    arc_sim *s = const_cast<arc_sim *>(cs);
    if (i < 0 || i >= s->n)
        return -1;
    int v = 0;
    HIPCHK_RET(s, hipMemcpyAsync(&v, s->dev.err + i, sizeof(int), hipMemcpyDeviceToHost, s->stream), "arcx_world_error", -1);
    HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arcx_world_error", -1);
    return v;
}

int arcx_last_error(const arc_sim *s) { return s->err; }
int arcx_fast_divisors(const arc_sim *s)
{
    const Model &M = s->su.model;
    return M.terr.fast_cell ? 1 : 0;
}
const char *arcx_error_message(const arc_sim *s) { return s->err ? s->msg : ""; }
void arcx_clear_error(arc_sim *s) { s->err = 0; s->msg[0] = 0; }

/* host snapshot of world i in the reference layout */
const arc_world *arc_sim_world(const arc_sim *cs, int i)
{
    arc_sim *s = const_cast<arc_sim *>(cs);
    if (i < 0 || i >= s->n)
        return nullptr;
    if (!s->mirror) {
        s->mirror = (arc_world *)calloc(1, sizeof(arc_world));
        if (!s->mirror)
            return nullptr;
    }
    std::vector<double> st(arcx_state_size()), ter(s->dev.cells + 1);
    if (arcx_state(s, i, st.data()) || arcx_terrain(s, i, ter.data()))
        return nullptr;
    arc_world *w = s->mirror;
    memset(w, 0, sizeof(*w));
    const HostWorld &T = s->su.tmpl;
    w->gravity.x = T.g.x; w->gravity.y = T.g.y; w->gravity.z = T.g.z;
    w->mu_static = T.mus;
    w->mu_dynamic = T.mud;
    w->n_bodies = T.nb;
    w->n_joints = T.nj;
    w->n_spheres = T.ns;
    w->n_visuals = robot_visuals(w->visuals, ARC_MAX_VISUALS);
    /* the contacts of the last substep of the world's last step, in the
     * reference's record order */
    {
        std::vector<float> c(1 + 7 * MAXC);
        HIPCHK_RET(s, launch_dump_contacts(s->dev, s->d_model, i, s->contact_layout, s->d_dump, s->stream),
                   "arc_sim_world", nullptr);
        HIPCHK_RET(s, hipMemcpyAsync(c.data(), s->d_dump, c.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream),
                   "arc_sim_world", nullptr);
        HIPCHK_RET(s, hipStreamSynchronize(s->stream), "arc_sim_world", nullptr);
        w->n_contacts = (int)c[0];
        for (int k = 0; k < w->n_contacts; ++k) {
            const float *r = &c[1 + 7 * k];
            arc_contact &C = w->contacts[k];
            C.body = (int)r[0];
            C.local.x = r[1]; C.local.y = r[2]; C.local.z = r[3];
            C.n.x = r[4]; C.n.y = r[5]; C.n.z = r[6];
        }
    }
    int o = 0;
    for (int b = 0; b < NB; ++b) {
        arc_body &B = w->bodies[b];
        Body hb;
        memset(&hb, 0, sizeof(hb));
        hb.x = v3((float)st[o], (float)st[o + 1], (float)st[o + 2]);
        hb.q = q4((float)st[o + 3], (float)st[o + 4], (float)st[o + 5], (float)st[o + 6]);
        B.x.x = st[o]; B.x.y = st[o + 1]; B.x.z = st[o + 2];
        B.q.w = st[o + 3]; B.q.x = st[o + 4]; B.q.y = st[o + 5]; B.q.z = st[o + 6];
        B.v.x = st[o + 7]; B.v.y = st[o + 8]; B.v.z = st[o + 9];
        B.w.x = st[o + 10]; B.w.y = st[o + 11]; B.w.z = st[o + 12];
        o += 13;
        B.x_prev = B.x;
        B.q_prev = B.q;
        const BodyK &K = T.bk[b];
        B.mass = K.mass; B.inv_mass = K.invm;
        B.inertia.x = K.I.x; B.inertia.y = K.I.y; B.inertia.z = K.I.z;
        B.inv_inertia.x = K.invI.x; B.inv_inertia.y = K.invI.y; B.inv_inertia.z = K.invI.z;
        bsync<KGEN>(hb, K);
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) {
                B.rot.m[a][c] = hb.R.e[a][c];
                B.inv_iw.m[a][c] = hb.Iw.e[a][c];
            }
    }
    for (int j = 0; j < NJ; ++j) {
        arc_joint &J = w->joints[j];
        const HostJoint &H = T.j[j];
        const JointK &K = s->su.model.joint[j];
        J.parent = H.parent; J.child = H.child;
        J.r1.x = H.r1.x; J.r1.y = H.r1.y; J.r1.z = H.r1.z;
        J.r2.x = H.r2.x; J.r2.y = H.r2.y; J.r2.z = H.r2.z;
        J.a1.x = H.a1.x; J.a1.y = H.a1.y; J.a1.z = H.a1.z;
        J.a2.x = H.a2.x; J.a2.y = H.a2.y; J.a2.z = H.a2.z;
        J.b1.x = H.b1.x; J.b1.y = H.b1.y; J.b1.z = H.b1.z;
        J.b2.x = H.b2.x; J.b2.y = H.b2.y; J.b2.z = H.b2.z;
        J.q_rel0.w = H.qrel0.w; J.q_rel0.x = H.qrel0.x; J.q_rel0.y = H.qrel0.y; J.q_rel0.z = H.qrel0.z;
        J.lo = H.lo; J.hi = H.hi; J.kp = H.kp; J.kd = H.kd; J.max_torque = H.tmax;
        J.target = st[o]; J.angle = st[o + 1]; J.rate = st[o + 2]; J.torque = st[o + 3];
        o += 4;
        for (int c = 0; c < 4; ++c)
            J.lim[c] = K.lim[c];
    }
    for (int k = 0; k < NS; ++k) {
        arc_sphere &S = w->spheres[k];
        const HostSphere &H = T.s[k];
        S.body = H.body;
        S.pos.x = H.at.x; S.pos.y = H.at.y; S.pos.z = H.at.z;
        S.radius = H.rad;
        S.touch = (int)st[o]; S.anchored = (int)st[o + 1];
        S.anchor.x = st[o + 2]; S.anchor.y = st[o + 3]; S.anchor.z = st[o + 4];
        S.anchor_n.x = st[o + 5]; S.anchor_n.y = st[o + 6]; S.anchor_n.z = st[o + 7];
        o += 8;
    }
    w->unstable = (int)st[o++];
    w->time = st[o++];
    arc_terrain &t = w->terrain;
    t.nx = T.t.nx; t.ny = T.t.ny; t.cell = T.t.cell; t.ox = T.t.ox; t.oy = T.t.oy;
    for (int c = 0; c < s->dev.cells; ++c)
        t.h[c] = ter[c];
    t.outside = ter[s->dev.cells];
    return w;
}

} /* extern "C" */
