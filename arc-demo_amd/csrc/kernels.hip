/*
 * kernels.hip -- HIP kernels of the batched engine for gfx950.
 *
 *   (k_step, the lane-form step kernel, lives in step_kernel.h)
 *   k_reset         regenerates terrain and re-poses listed worlds, one
 *                   wavefront per world (csrc/arc_sim.c:88-121)
 *   k_observe / k_heights / k_terrain_at / k_dump   read-outs
 *   k_init_template every world set to the robot as built on flat ground
 */
#include "device_state.h"
#include "reset_engine.h"

namespace xpbd {

static inline size_t round64(size_t n) { return (n + 63) & ~(size_t)63; }

/* float counts per world in the SoA block (terrain and contact list apart) */
static const int SOA_FLOATS = NB * 3 + NB * 4 + NB * 3 + NB * 3 + NJ + 3 * NJ + NS * 3 + NS * 3 + 1 + 1 + 1;

static size_t tcells_of(int grid) { return (size_t)(grid + 2 * TPAD) * (grid + 2 * TPAD); }

size_t dev_bytes_per_world(int grid, int cslots)
{
    return sizeof(float) * ((size_t)SOA_FLOATS + 2 + (size_t)cslots * 6) + tcells_of(grid) + NS /* touch */
           + NCB /* contact counts */ + sizeof(uint32_t) + 2 * sizeof(int);
}

hipError_t dev_alloc(DevWorlds &d, int n, int grid, int cslots)
{
    d.n = n;
    d.stride = round64((size_t)(n > 0 ? n : 1));
    d.cells = grid * grid;
    d.tpitch = grid + 2 * TPAD;
    d.tcells = (int)tcells_of(grid);
    size_t S = d.stride;
    size_t floats = S * ((size_t)SOA_FLOATS + 2 + (size_t)cslots * 6);
    size_t bytes = floats * sizeof(float) + S * (size_t)d.tcells + S * NS + S * NCB + S * sizeof(uint32_t)
                   + 2 * S * sizeof(int) + 1024;
    char *p = nullptr;
    hipError_t e = hipMalloc((void **)&p, bytes);
    if (e != hipSuccess)
        return e;
    d.base = p;
    d.bytes = bytes;
    d.lcbuf = nullptr;
    d.lcdummy = nullptr;
    float *f = (float *)p;
    d.x = f; f += NB * 3 * S;
    d.q = f; f += NB * 4 * S;
    d.v = f; f += NB * 3 * S;
    d.w = f; f += NB * 3 * S;
    d.target = f; f += NJ * S;
    d.jout = f; f += 3 * NJ * S;
    d.anc = f; f += NS * 3 * S;
    d.ancn = f; f += NS * 3 * S;
    d.time = f; f += S;
    d.outside = f; f += S;
    d.difficulty = f; f += S;
    d.tbase = f; f += S;
    d.cbuf = f; f += (size_t)cslots * 6 * S;
    d.unstable = (int *)f;
    d.err = d.unstable + S;
    d.anchored = (uint32_t *)(d.err + S);
    d.touch = (uint8_t *)(d.anchored + S);
    d.ncontact = d.touch + S * NS;
    d.tq = (int8_t *)(d.ncontact + S * NCB);
    return hipSuccess;
}

void dev_free(DevWorlds &d)
{
    if (d.base)
        (void)hipFree(d.base);
    d.base = nullptr;
    d.x = nullptr;
}

/* ------------------------------------------------------------------ */


hipError_t launch_step(const DevWorlds &d, const Model *m_dev, const float *targets, const StepBatch *batch, int total,
                       int block, int tile, int grid_blocks, int *work, int leg, hipStream_t s)
{
    if (!batch && total <= 0)
        return hipSuccess;
    switch (block) {
    case 64: return launch_step_b64(d, m_dev, targets, batch, total, tile, grid_blocks, work, leg, s);
    case 128: return launch_step_b128(d, m_dev, targets, batch, total, tile, grid_blocks, work, leg, s);
    case 256: return launch_step_b256(d, m_dev, targets, batch, total, tile, grid_blocks, work, leg, s);
    default: return hipErrorInvalidValue;
    }
}

int step_resident_blocks(int block, int tile, int leg)
{
    switch (block) {
    case 64: return step_resident_blocks_b64(tile, leg);
    case 128: return step_resident_blocks_b128(tile, leg);
    case 256: return step_resident_blocks_b256(tile, leg);
    default: return 0;
    }
}

/* ------------------------------------------------------------------ */
/* reset: terrain generation, joint noise, yaw, ground scan, FK          */

__global__ __launch_bounds__(RESET_LANES) void k_reset(DevWorlds d, const Model *__restrict__ m,
                                                       const ResetK *__restrict__ rk, const int *__restrict__ ids,
                                                       const uint32_t *__restrict__ seeds,
                                                       const double *__restrict__ diff_d,
                                                       const float *__restrict__ diff_f, int n_ids,
                                                       const int *__restrict__ count_dev)
{
    __shared__ float hbuf[GRID_LIMIT_CELLS];
    __shared__ int8_t qbuf[(64 + 2 * TPAD) * (64 + 2 * TPAD)];
    __shared__ float lat[LAT_LIMIT];
    __shared__ float draws[NJ + 1];
    __shared__ float gpart[RESET_LANES / 64];
    __shared__ float s_ang[NJ];
    __shared__ Q4 s_rel[NJ];  /* qaxis(a1, angle) * qrel0 of each joint */
    __shared__ float s_pose[NB * 7];
    const Model &M = *m;
    const ResetK &R = *rk;
    const int tid = threadIdx.x;
    const int count = count_dev ? *count_dev : n_ids;
    for (int k = blockIdx.x; k < count; k += gridDim.x) {
    const int i = ids[k];
    const float dfc = diff_d ? (float)diff_d[k] : diff_f[k];
    const int n = R.grid, cells = n * n;
    const bool gen = dfc > 0.0f;
    const size_t S = d.stride;
    const int pitch = d.tpitch;

    /* the draws: one stream, cut into RESET_LANES contiguous chunks that the
     * lanes reach by jumping ahead (without terrain: 13 draws, lane 0) */
    if (gen)
        reset_draws_lane(seeds[k], gen, n, R.jtab, R.chunk, tid, lat, hbuf, draws);
    else if (tid == 0)
        reset_draws(seeds[k], gen, n, lat, hbuf, draws);
    __syncthreads();
    /* quantisation steps of the columns (csrc/arc_terrain.c:54-71) */
    for (int c = tid; c < cells; c += RESET_LANES)
        hbuf[c] = gen ? column_step(lat, n, dfc, hbuf[c], c) : 0.0f;
    /* each joint's rotation from its angle (csrc/arc_sim.c:92-100), in
     * parallel: the chain below then only multiplies */
    if (tid < NJ) {
        const JointK &J = M.joint[tid];
        const float ang = clampf_(R.def[tid] + draws[tid], R.lo[tid], R.hi[tid]);
        s_ang[tid] = ang;
        s_rel[tid] = qmul(qaxis(J.a1, ang), J.qrel0);
    }
    __syncthreads();
    /* base: the pre-shift height of the column at the origin */
    float base = 0.0f;
    if (gen) {
        int ix = (int)floorf((0.0f - M.terr.ox) / M.terr.cell), iy = (int)floorf((0.0f - M.terr.oy) / M.terr.cell);
        base = 0.03f * hbuf[ix + iy * n];
    }
    int8_t *gq = d.tq + (size_t)i * d.tcells;
    for (int c = tid; c < d.tcells; c += RESET_LANES) {
        int ix = c % pitch - TPAD, iy = c / pitch - TPAD;
        int8_t v = (ix >= 0 && iy >= 0 && ix < n && iy < n) ? (int8_t)(int)hbuf[ix + iy * n] : (int8_t)Q_OUT;
        qbuf[c] = v;
        gq[c] = v;
    }
    __syncthreads();
    TerrainRef tl;
    tl.q = qbuf + TPAD + TPAD * pitch;
    tl.pitch = pitch;
    tl.base = base;
    tl.outside = -100.0f;

    /* the footprint scan split over the lanes, then the maximum */
    float ground = ground_part(M.terr, tl, R.footprint, R.cell, tid, RESET_LANES);
    for (int off = 32; off; off >>= 1)
        ground = maxf_(ground, __shfl_xor(ground, off, 64));
    if ((tid & 63) == 0)
        gpart[tid >> 6] = ground;
    __syncthreads();
    /* forward kinematics (pose_bodies): the torso, then each leg's chain of
     * three joints in its own lane */
    if (tid < LEGS) {
        float g = gpart[0];
        for (int w = 1; w < RESET_LANES / 64; ++w)
            g = maxf_(g, gpart[w]);
        V3 px = v3(0.0f, 0.0f, g + R.stand + 0.03f);
        Q4 pq = qunit(qaxis(v3(0.0f, 0.0f, 1.0f), draws[NJ]));
        if (tid == 0) {
            s_pose[0] = px.x; s_pose[1] = px.y; s_pose[2] = px.z;
            s_pose[3] = pq.w; s_pose[4] = pq.x; s_pose[5] = pq.y; s_pose[6] = pq.z;
        }
        for (int jj = 0; jj < 3; ++jj) {
            const int j = 3 * tid + jj, c = joint_child(j);
            const JointK &J = M.joint[j];
            V3 anchor = add(px, qrot(pq, J.r1));
            Q4 cq = qunit(qmul(pq, s_rel[j]));
            V3 cx = sub(anchor, qrot(cq, J.r2));
            float *o = s_pose + 7 * c;
            o[0] = cx.x; o[1] = cx.y; o[2] = cx.z;
            o[3] = cq.w; o[4] = cq.x; o[5] = cq.y; o[6] = cq.z;
            px = cx;
            pq = cq;
        }
    }
    __syncthreads();
    /* the world's state, one value per lane */
    for (int e = tid; e < NB * 13; e += RESET_LANES) {
        const int b = e / 13, f = e % 13;
        if (f < 3)
            d.x[(b * 3 + f) * S + i] = s_pose[7 * b + f];
        else if (f < 7)
            d.q[(b * 4 + f - 3) * S + i] = s_pose[7 * b + f];
        else if (f < 10)
            d.v[(b * 3 + f - 7) * S + i] = 0.0f;
        else
            d.w[(b * 3 + f - 10) * S + i] = 0.0f;
    }
    for (int e = tid; e < NJ * 4; e += RESET_LANES) {
        const int j = e >> 2, f = e & 3;
        if (f == 0)
            d.target[j * S + i] = s_ang[j];
        else
            d.jout[((f - 1) * NJ + j) * S + i] = f == 1 ? s_ang[j] : 0.0f;
    }
    for (int e = tid; e < NS * 7; e += RESET_LANES) {
        const int sp = e / 7, f = e % 7;
        if (f == 0)
            d.touch[sp * S + i] = 0;
        else if (f < 4)
            d.anc[(sp * 3 + f - 1) * S + i] = 0.0f;
        else
            d.ancn[(sp * 3 + f - 4) * S + i] = 0.0f;
    }
    if (tid < NCB)
        d.ncontact[(size_t)tid * S + i] = 0; /* no contacts after a reset (csrc/arc_phys.c:780) */
    if (tid == 0) {
        d.anchored[i] = 0;
        d.unstable[i] = 0;
        d.time[i] = 0.0f;
        d.outside[i] = -100.0f;
        d.tbase[i] = base;
        d.difficulty[i] = dfc;
        d.err[i] = 0;
    }
    __syncthreads(); /* LDS is reused by the next entry */
    }
}

hipError_t launch_reset(const DevWorlds &d, const Model *m_dev, const ResetK *r_dev, const int *ids,
                        const uint32_t *seeds, const double *difficulty_d, const float *difficulty_f, int n_ids,
                        const int *count_dev, int max_blocks, hipStream_t s)
{
    int blocks = count_dev ? max_blocks : (n_ids < max_blocks || max_blocks <= 0 ? n_ids : max_blocks);
    if (blocks <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_reset, dim3(blocks), dim3(RESET_LANES), 0, s, d, m_dev, r_dev, ids, seeds, difficulty_d, difficulty_f,
                       n_ids, count_dev);
    return hipGetLastError();
}

/* fall rules of the task after a step, arc/task.py:180-203; comparisons in
 * binary64 on the binary32 state, as the task makes them on float64 copies */
__global__ void k_falls(DevWorlds d, const Model *__restrict__ m, TaskK tk, TaskState ts,
                        const StepBatch *__restrict__ batch)
{
    const size_t S = d.stride;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (batch ? batch->total : d.n))
        return;
    const int i = batch ? batch_world(batch, batch->n_ranges, t) : t;
    float r[13];
    for (int c = 0; c < 3; ++c) r[c] = d.x[c * S + i];
    for (int c = 0; c < 4; ++c) r[3 + c] = d.q[c * S + i];
    for (int c = 0; c < 3; ++c) r[7 + c] = d.v[c * S + i];
    for (int c = 0; c < 3; ++c) r[10 + c] = d.w[c * S + i];
    bool bad = d.unstable[i] != 0;
    for (int c = 0; c < 13; ++c)
        bad |= !isfinite(r[c]);
    float torso = 0.0f;
    static_for<0, NS>([&](auto s) {
        if (sphere_channel(s) == LEGS)
            torso = maxf_(torso, (float)d.touch[s * S + i] * m->inv_sub);
    });
    int steps = torso > 0.0f ? ts.torso_steps[i] + 1 : 0;
    M3 R;
    m_from_q(R, q4(r[3], r[4], r[5], r[6]));
    bool vd = fabs((double)r[0]) > tk.half_extent || fabs((double)r[1]) > tk.half_extent || (double)r[2] < tk.void_z;
    bool flip = (double)R.e[2][2] < tk.flip_up_z;
    bool beached = steps > tk.beached_steps;
    int cause = bad ? 1 : vd ? 2 : flip ? 3 : beached ? 4 : 0;
    ts.cause[i] = cause;
    if (cause) {
        uint32_t ep = ts.episode[i] + 1;
        ts.episode[i] = ep;
        int slot = atomicAdd(ts.count, 1);
        ts.list[slot] = i;
        ts.seeds[slot] = (uint32_t)i + (uint32_t)tk.n * ep;
        ts.diff[slot] = d.difficulty[i];
        steps = 0;
        atomicAdd(&ts.stats[cause], 1ull);
    }
    ts.torso_steps[i] = steps;
}

__global__ void k_fill_int(int *p, int v, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

hipError_t launch_fill_int(int *p, int v, int n, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t launch_falls(const DevWorlds &d, const Model *m_dev, const TaskK &tk, const TaskState &ts,
                        const StepBatch *batch, hipStream_t s)
{
    if (d.n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_falls, dim3((d.n + 255) / 256), dim3(256), 0, s, d, m_dev, tk, ts, batch);
    return hipGetLastError();
}

/* ------------------------------------------------------------------ */

__global__ void k_init_template(DevWorlds d, const TemplateK *__restrict__ tk, float outside)
{
    const size_t S = d.stride;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)d.n; i += (size_t)gridDim.x * blockDim.x) {
        for (int b = 0; b < NB; ++b) {
            d.x[(b * 3 + 0) * S + i] = tk->x[b].x; d.x[(b * 3 + 1) * S + i] = tk->x[b].y; d.x[(b * 3 + 2) * S + i] = tk->x[b].z;
            d.q[(b * 4 + 0) * S + i] = tk->q[b].w; d.q[(b * 4 + 1) * S + i] = tk->q[b].x;
            d.q[(b * 4 + 2) * S + i] = tk->q[b].y; d.q[(b * 4 + 3) * S + i] = tk->q[b].z;
            for (int c = 0; c < 3; ++c) {
                d.v[(b * 3 + c) * S + i] = 0.0f;
                d.w[(b * 3 + c) * S + i] = 0.0f;
            }
        }
        for (int j = 0; j < NJ; ++j) {
            d.target[j * S + i] = tk->target[j];
            d.jout[(0 * NJ + j) * S + i] = tk->angle[j];
            d.jout[(1 * NJ + j) * S + i] = 0.0f;
            d.jout[(2 * NJ + j) * S + i] = 0.0f;
        }
        for (int s = 0; s < NS; ++s) {
            d.touch[s * S + i] = 0;
            for (int c = 0; c < 3; ++c) {
                d.anc[(s * 3 + c) * S + i] = 0.0f;
                d.ancn[(s * 3 + c) * S + i] = 0.0f;
            }
        }
        for (int c = 0; c < NCB; ++c)
            d.ncontact[(size_t)c * S + i] = 0;
        d.anchored[i] = 0;
        d.unstable[i] = 0;
        d.time[i] = 0.0f;
        d.outside[i] = outside;
        d.tbase[i] = 0.0f;
        d.difficulty[i] = 0.0f;
        d.err[i] = 0;
    }
}

/* flat template terrain: steps 0 inside the grid, the border outside */
__global__ void k_init_terrain(DevWorlds d)
{
    const int n = d.tpitch - 2 * TPAD;
    size_t total = (size_t)d.n * d.tcells;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        int c = (int)(t % d.tcells);
        int ix = c % d.tpitch - TPAD, iy = c / d.tpitch - TPAD;
        d.tq[t] = (ix >= 0 && iy >= 0 && ix < n && iy < n) ? (int8_t)0 : (int8_t)Q_OUT;
    }
}

hipError_t launch_init_template(const DevWorlds &d, const TemplateK *tk_dev, float outside, hipStream_t s)
{
    int blocks = (d.n + 255) / 256;
    if (blocks < 1)
        blocks = 1;
    hipLaunchKernelGGL(k_init_template, dim3(blocks), dim3(256), 0, s, d, tk_dev, outside);
    hipLaunchKernelGGL(k_init_terrain, dim3(4096), dim3(256), 0, s, d);
    return hipGetLastError();
}

/* ------------------------------------------------------------------ */
/* read-outs, csrc/arc_sim.c:140-210                                    */

__global__ void k_observe(DevWorlds d, const Model *__restrict__ m, double *root, double *rot, double *joints,
                          double *contacts, int *flags)
{
    const Model &M = *m;
    const size_t S = d.stride;
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)d.n)
        return;
    Q4 q = q4(d.q[0 * S + i], d.q[1 * S + i], d.q[2 * S + i], d.q[3 * S + i]);
    if (root) {
        double *r = root + i * 13;
        r[0] = d.x[0 * S + i]; r[1] = d.x[1 * S + i]; r[2] = d.x[2 * S + i];
        r[3] = q.w; r[4] = q.x; r[5] = q.y; r[6] = q.z;
        r[7] = d.v[0 * S + i]; r[8] = d.v[1 * S + i]; r[9] = d.v[2 * S + i];
        r[10] = d.w[0 * S + i]; r[11] = d.w[1 * S + i]; r[12] = d.w[2 * S + i];
    }
    if (rot) {
        M3 R;
        m_from_q(R, q);
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c)
                rot[i * 9 + a * 3 + c] = R.e[a][c];
    }
    if (joints) {
        for (int j = 0; j < 3 * NJ; ++j)
            joints[i * 36 + j] = d.jout[j * S + i];
    }
    if (contacts) {
        float o[NCH];
        for (int c = 0; c < NCH; ++c)
            o[c] = 0.0f;
        static_for<0, NS>([&](auto s) {
            constexpr int ch = sphere_channel(s);
            o[ch] = maxf_(o[ch], (float)d.touch[s * S + i] * M.inv_sub);
        });
        for (int c = 0; c < NCH; ++c)
            contacts[i * NCH + c] = o[c];
    }
    if (flags)
        flags[i] = d.unstable[i];
}

hipError_t launch_observe(const DevWorlds &d, const Model *m_dev, double *root, double *rot, double *joints,
                          double *contacts, int *flags, hipStream_t s)
{
    if (d.n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_observe, dim3((d.n + 255) / 256), dim3(256), 0, s, d, m_dev, root, rot, joints, contacts, flags);
    return hipGetLastError();
}

__global__ void k_heights(DevWorlds d, const Model *__restrict__ m, const double *__restrict__ offsets, int k, double *out)
{
    const Model &M = *m;
    const size_t S = d.stride;
    size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)d.n * k)
        return;
    size_t i = t / k;
    int j = (int)(t % k);
    Q4 q = q4(d.q[0 * S + i], d.q[1 * S + i], d.q[2 * S + i], d.q[3 * S + i]);
    V3 fwd = qrot(q, v3(1.0f, 0.0f, 0.0f));
    float yaw = f_atan2(fwd.y, fwd.x), c = f_cos(yaw), sn = f_sin(yaw);
    float f = (float)offsets[2 * j], l = (float)offsets[2 * j + 1];
    float px = d.x[0 * S + i], py = d.x[1 * S + i];
    TerrainRef tr = terrain_ref(d, i);
    tr.base = d.tbase[i];
    tr.outside = d.outside[i];
    out[t] = terrain_height(M.terr, tr, px + c * f - sn * l, py + sn * f + c * l);
}

hipError_t launch_heights(const DevWorlds &d, const Model *m_dev, const double *offsets, int k, double *out,
                          hipStream_t s)
{
    size_t total = (size_t)d.n * k;
    if (total == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_heights, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, d, m_dev, offsets, k, out);
    return hipGetLastError();
}

__global__ void k_terrain_at(DevWorlds d, const Model *__restrict__ m, const int *__restrict__ ids,
                             const double *__restrict__ xy, int mcount, double *out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= mcount)
        return;
    int i = ids[t];
    TerrainRef tr = terrain_ref(d, i);
    tr.base = d.tbase[i];
    tr.outside = d.outside[i];
    out[t] = terrain_height(m->terr, tr, (float)xy[2 * t], (float)xy[2 * t + 1]);
}

hipError_t launch_terrain_at(const DevWorlds &d, const Model *m_dev, const int *ids, const double *xy, int mcount,
                             double *out, hipStream_t s)
{
    if (mcount <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_terrain_at, dim3((mcount + 255) / 256), dim3(256), 0, s, d, m_dev, ids, xy, mcount, out);
    return hipGetLastError();
}

/* world i in the oracle's dump layout: bodies (x3 q4 v3 w3), joints (target
 * angle rate torque), spheres (touch anchored anchor3 anchor_n3), unstable, time */
__global__ void k_dump(DevWorlds d, int i, float *out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    const size_t S = d.stride;
    int o = 0;
    for (int b = 0; b < NB; ++b) {
        for (int c = 0; c < 3; ++c) out[o++] = d.x[(b * 3 + c) * S + i];
        for (int c = 0; c < 4; ++c) out[o++] = d.q[(b * 4 + c) * S + i];
        for (int c = 0; c < 3; ++c) out[o++] = d.v[(b * 3 + c) * S + i];
        for (int c = 0; c < 3; ++c) out[o++] = d.w[(b * 3 + c) * S + i];
    }
    for (int j = 0; j < NJ; ++j) {
        out[o++] = d.target[j * S + i];
        out[o++] = d.jout[(0 * NJ + j) * S + i];
        out[o++] = d.jout[(1 * NJ + j) * S + i];
        out[o++] = d.jout[(2 * NJ + j) * S + i];
    }
    for (int s = 0; s < NS; ++s) {
        out[o++] = (float)d.touch[s * S + i];
        out[o++] = (float)((d.anchored[i] >> s) & 1u);
        for (int c = 0; c < 3; ++c) out[o++] = d.anc[(s * 3 + c) * S + i];
        for (int c = 0; c < 3; ++c) out[o++] = d.ancn[(s * 3 + c) * S + i];
    }
    out[o++] = (float)d.unstable[i];
    out[o++] = d.time[i];
}

/* the contacts world i recorded in the last substep of its last step, in
 * the reference's record order (spheres in index order: the torso's, then
 * each leg's femur sphere and tibia spheres, csrc/arc_phys.c:488-500):
 * out[0] = count, then per contact body index, local point (3), normal (3).
 * layout: where the step left its lists (1: cbuf, slots of the contact
 * body's range; 2 / 3: lcbuf, four / two lanes per world, every lane's list
 * holding the torso's and its own legs' contacts). */
__global__ void k_dump_contacts(DevWorlds d, const Model *__restrict__ m, int i, int layout, float *out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    const Model &M = *m;
    const size_t S = d.stride;
    int o = 1, total = 0;
    for (int e = 0; e < NCB; ++e) {
        /* torso, then leg g's femur (contact body 1 + g) and tibia (5 + g) */
        const int g = (e - 1) / 2;
        const int ci = e == 0 ? 0 : ((e - 1) % 2 == 0 ? 1 + g : 5 + g);
        const int kind = ci == 0 ? 0 : ci <= LEGS ? 1 : 2;
        const int n = layout ? d.ncontact[(size_t)ci * S + i] : 0;
        const float *base;
        size_t stride, w;
        int slot0;
        if (layout == 1) {
            base = d.cbuf;
            stride = S;
            w = (size_t)i;
            slot0 = M.cslot[ci];
        } else {
            const int K = layout == 2 ? 1 : 2, P = LEGS / K, q = kind == 0 ? 0 : g / K, l = kind == 0 ? 0 : g % K;
            base = d.lcbuf;
            stride = (size_t)P * S;
            w = (size_t)i * P + q;
            slot0 = kind == 0 ? 0 : kind == 1 ? M.ccap[0] + l * M.ccap[1] : M.ccap[0] + K * M.ccap[1] + l * M.ccap[5];
        }
        for (int k = 0; k < n && total < MAXC; ++k, ++total) {
            const float *p = base + (size_t)(slot0 + k) * 6 * stride + w;
            out[o++] = (float)cbody(ci);
            for (int f = 0; f < 6; ++f)
                out[o++] = p[f * stride];
        }
    }
    out[0] = (float)total;
}

hipError_t launch_dump_contacts(const DevWorlds &d, const Model *m_dev, int i, int layout, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_dump_contacts, dim3(1), dim3(64), 0, s, d, m_dev, i, layout, out);
    return hipGetLastError();
}

hipError_t launch_dump(const DevWorlds &d, int i, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_dump, dim3(1), dim3(64), 0, s, d, i, out);
    return hipGetLastError();
}

} /* namespace xpbd */
