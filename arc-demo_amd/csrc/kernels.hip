/*
 * kernels.hip -- HIP kernels of the batched engine for gfx950.
 *
 *   k_step          one control step per world, one world per lane
 *                   (lane_engine.h); a grid-stride loop over a list of world
 *                   ranges, so one launch serves uneven per-request batches
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
           + sizeof(uint32_t) + 2 * sizeof(int);
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
    size_t bytes = floats * sizeof(float) + S * (size_t)d.tcells + S * NS + S * sizeof(uint32_t) + 2 * S * sizeof(int) + 1024;
    char *p = nullptr;
    hipError_t e = hipMalloc((void **)&p, bytes);
    if (e != hipSuccess)
        return e;
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
    d.tq = (int8_t *)(d.touch + S * NS);
    return hipSuccess;
}

void dev_free(DevWorlds &d)
{
    if (d.x)
        (void)hipFree(d.x);
    d.x = nullptr;
}

/* ------------------------------------------------------------------ */

__device__ __forceinline__ void load_world(const DevWorlds &d, size_t i, LaneWorld &W)
{
    const size_t S = d.stride;
    static_for<0, NB>([&](auto b) {
        W.b[b].x = v3(d.x[(b * 3 + 0) * S + i], d.x[(b * 3 + 1) * S + i], d.x[(b * 3 + 2) * S + i]);
        W.b[b].q = q4(d.q[(b * 4 + 0) * S + i], d.q[(b * 4 + 1) * S + i], d.q[(b * 4 + 2) * S + i],
                      d.q[(b * 4 + 3) * S + i]);
        W.b[b].v = v3(d.v[(b * 3 + 0) * S + i], d.v[(b * 3 + 1) * S + i], d.v[(b * 3 + 2) * S + i]);
        W.b[b].w = v3(d.w[(b * 3 + 0) * S + i], d.w[(b * 3 + 1) * S + i], d.w[(b * 3 + 2) * S + i]);
    });
    W.anchored = d.anchored[i];
    W.unstable = d.unstable[i];
    W.time = d.time[i];
}

__device__ __forceinline__ void store_world(const DevWorlds &d, size_t i, const LaneWorld &W)
{
    const size_t S = d.stride;
    static_for<0, NB>([&](auto b) {
        d.x[(b * 3 + 0) * S + i] = W.b[b].x.x; d.x[(b * 3 + 1) * S + i] = W.b[b].x.y; d.x[(b * 3 + 2) * S + i] = W.b[b].x.z;
        d.q[(b * 4 + 0) * S + i] = W.b[b].q.w; d.q[(b * 4 + 1) * S + i] = W.b[b].q.x;
        d.q[(b * 4 + 2) * S + i] = W.b[b].q.y; d.q[(b * 4 + 3) * S + i] = W.b[b].q.z;
        d.v[(b * 3 + 0) * S + i] = W.b[b].v.x; d.v[(b * 3 + 1) * S + i] = W.b[b].v.y; d.v[(b * 3 + 2) * S + i] = W.b[b].v.z;
        d.w[(b * 3 + 0) * S + i] = W.b[b].w.x; d.w[(b * 3 + 1) * S + i] = W.b[b].w.y; d.w[(b * 3 + 2) * S + i] = W.b[b].w.z;
    });
    static_for<0, NJ>([&](auto j) {
        d.jout[(0 * NJ + j) * S + i] = W.angle[j];
        d.jout[(1 * NJ + j) * S + i] = W.rate[j];
        d.jout[(2 * NJ + j) * S + i] = W.torque[j];
    });
    d.anchored[i] = W.anchored;
    static_for<0, NS>([&](auto s) { d.touch[s * S + i] = (uint8_t)touch_of(W, s); });
    d.unstable[i] = W.unstable;
    d.time[i] = W.time;
}

/* friction anchor of sphere s of world i, read and written in its HBM
 * arrays only while the sphere is in contact */
struct GlobalAnchor {
    float *a, *n;
    size_t S, i;
    int s;
    __device__ __forceinline__ V3 get(const float *p) const
    {
        return v3(p[(s * 3 + 0) * S + i], p[(s * 3 + 1) * S + i], p[(s * 3 + 2) * S + i]);
    }
    __device__ __forceinline__ void put(float *p, V3 v)
    {
        p[(s * 3 + 0) * S + i] = v.x;
        p[(s * 3 + 1) * S + i] = v.y;
        p[(s * 3 + 2) * S + i] = v.z;
    }
    __device__ __forceinline__ V3 point() const { return get(a); }
    __device__ __forceinline__ V3 normal() const { return get(n); }
    __device__ __forceinline__ void set_point(V3 v) { put(a, v); }
    __device__ __forceinline__ void set_normal(V3 v) { put(n, v); }
};
struct GlobalAnchors {
    float *a, *n;
    size_t S, i;
    __device__ __forceinline__ GlobalAnchor at(int s) const { return GlobalAnchor{ a, n, S, i, s }; }
    __device__ __forceinline__ GlobalAnchors fresh() const { return GlobalAnchors{ a, n, S, launder(i) }; }
};

/* substep-start poses of the block's worlds in LDS: [body * 7 + field][thread] */
template <int BLOCK> struct LdsPoses {
    float *p;
    int tid;
    __device__ __forceinline__ void store(int b, const Pose0 &v)
    {
        float *q = p + b * 7 * BLOCK + tid;
        q[0] = v.x0.x; q[BLOCK] = v.x0.y; q[2 * BLOCK] = v.x0.z;
        q[3 * BLOCK] = v.q0.w; q[4 * BLOCK] = v.q0.x; q[5 * BLOCK] = v.q0.y; q[6 * BLOCK] = v.q0.z;
    }
    __device__ __forceinline__ Pose0 load(int b) const
    {
        const float *q = p + b * 7 * BLOCK + tid;
        Pose0 v;
        v.x0 = v3(q[0], q[BLOCK], q[2 * BLOCK]);
        v.q0 = q4(q[3 * BLOCK], q[4 * BLOCK], q[5 * BLOCK], q[6 * BLOCK]);
        return v;
    }
};

/* The columns around a world's torso, copied into LDS once per control step:
 * TILE x TILE int8 heights of the padded grid, [column][thread].  Every lookup
 * inside the tile reads LDS; a lookup outside it (a leg stretched beyond the
 * tile, or a world that left the grid) reads the grid in HBM, which holds the
 * same bytes, so the heights are identical either way. */
constexpr int TILE = 16;
template <int BLOCK> struct LdsTileCells {
    const TerrainK &t;
    const TerrainRef &r;
    const int8_t *lt;
    int tx, ty;
    __device__ __forceinline__ float operator()(int ix, int iy) const
    {
        unsigned u = (unsigned)(ix - tx), v = (unsigned)(iy - ty);
        if (u < (unsigned)TILE && v < (unsigned)TILE)
            return dequant(r, lt[(u + v * TILE) * BLOCK]);
        return terrain_cell(t, r, ix, iy);
    }
};
template <int BLOCK> struct LdsTile {
    const int8_t *lt;
    int tx, ty;
    __device__ __forceinline__ LdsTileCells<BLOCK> cells(const TerrainK &t, const TerrainRef &r) const
    {
        return LdsTileCells<BLOCK>{ t, r, lt, tx, ty };
    }
};

/* fill this thread's tile of world i around the torso position (x, y) */
template <int BLOCK>
__device__ __forceinline__ LdsTile<BLOCK> fill_tile(int8_t *s_tile, const DevWorlds &d, const TerrainK &t, size_t i,
                                                    float x, float y)
{
    LdsTile<BLOCK> T;
    int8_t *lt = s_tile + threadIdx.x;
    T.lt = lt;
    const int n = d.tpitch - 2 * TPAD;
    if (d.tpitch < TILE) {
        T.tx = T.ty = 1 << 28; /* grid smaller than a tile: every lookup goes to HBM */
        return T;
    }
    /* clamped in float first: a non-finite torso still yields a valid tile */
    float fx = fminf(fmaxf((x - t.ox) / t.cell, -1.0e4f), 1.0e4f);
    float fy = fminf(fmaxf((y - t.oy) / t.cell, -1.0e4f), 1.0e4f);
    T.tx = min(max((int)floorf(fx) - TILE / 2, -TPAD), n + TPAD - TILE);
    T.ty = min(max((int)floorf(fy) - TILE / 2, -TPAD), n + TPAD - TILE);
    const int8_t *src = d.tq + i * d.tcells + (T.tx + TPAD) + (size_t)(T.ty + TPAD) * d.tpitch;
#pragma unroll 2
    for (int v = 0; v < TILE; ++v) {
#pragma unroll
        for (int w = 0; w < TILE / 4; ++w) {
            uint32_t word;
            __builtin_memcpy(&word, src + v * d.tpitch + 4 * w, 4);
#pragma unroll
            for (int b = 0; b < 4; ++b)
                lt[(4 * w + b + v * TILE) * BLOCK] = (int8_t)(word >> (8 * b));
        }
    }
    return T;
}

/* An index the optimiser cannot see through.  The world's SoA addresses are
 * recomputed from it for the stores at the end of the step instead of being
 * shared with the loads at its start, which would keep ~450 64-bit addresses
 * live across the whole step. */
__device__ __forceinline__ size_t opaque_index(size_t i)
{
    asm volatile("" : "+v"(i));
    return i;
}

template <class TT, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_step(DevWorlds d, const Model *__restrict__ m, const TT *__restrict__ targets,
                                                const Range *__restrict__ ranges, int n_ranges, int total)
{
    __shared__ float s_pose[NB * 7 * BLOCK];
    __shared__ int8_t s_tile[TILE * TILE * BLOCK];
    const Model &M = *m;
    LdsPoses<BLOCK> P0;
    P0.p = s_pose;
    P0.tid = threadIdx.x;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        int i = t;
        if (ranges) {
            int r = 0;
            while (r + 1 < n_ranges && ranges[r + 1].prefix <= t)
                ++r;
            i = ranges[r].start + (t - ranges[r].prefix);
        }
        const size_t S = d.stride;
        LaneWorld W;
        /* targets clamped to the limits, csrc/arc_sim.c:132-134 */
        static_for<0, NJ>([&](auto j) {
            float tg = (float)targets[(size_t)i * NJ + j];
            W.target[j] = clampf_(tg, M.joint[j].lo, M.joint[j].hi);
            d.target[j * S + i] = W.target[j];
        });
        if (d.unstable[i])
            continue; /* frozen until reset, csrc/arc_sim.c:135-136 */
        const LdsTile<BLOCK> tile = fill_tile<BLOCK>(s_tile, d, M.terr, (size_t)i, d.x[0 * S + i], d.x[1 * S + i]);
        load_world(d, i, W);
        GlobalAnchors A{ d.anc, d.ancn, S, (size_t)i };
        TerrainRef tr = terrain_ref(d, i);
        tr.base = d.tbase[i];
        tr.outside = d.outside[i];
        ContactStore cs;
        cs.base = d.cbuf;
        cs.stride = S;
        cs.world = i;
        int ov = lane_step(W, A, P0, M, tr, cs, tile);
        const size_t io = opaque_index((size_t)i);
        store_world(d, io, W);
        if (ov)
            d.err[io] |= ov;
    }
}

hipError_t launch_step(const DevWorlds &d, const Model *m_dev, const double *targets_d, const float *targets_f,
                       const Range *ranges, int n_ranges, int total, int block, int grid_blocks, hipStream_t s)
{
    if (total <= 0)
        return hipSuccess;
    if (block <= 0)
        block = 64;
    if (grid_blocks <= 0)
        grid_blocks = (total + block - 1) / block;
#define ARC_STEP_LAUNCH(B)                                                                                     \
    case B:                                                                                                    \
        if (targets_d)                                                                                         \
            hipLaunchKernelGGL((k_step<double, B>), dim3(grid_blocks), dim3(B), 0, s, d, m_dev, targets_d, ranges,   \
                               n_ranges, total);                                                               \
        else                                                                                                   \
            hipLaunchKernelGGL((k_step<float, B>), dim3(grid_blocks), dim3(B), 0, s, d, m_dev, targets_f, ranges,    \
                               n_ranges, total);                                                               \
        break;
    switch (block) {
        ARC_STEP_LAUNCH(64)
        ARC_STEP_LAUNCH(128)
        ARC_STEP_LAUNCH(256)
    default:
        return hipErrorInvalidValue;
    }
#undef ARC_STEP_LAUNCH
    return hipGetLastError();
}

/* ------------------------------------------------------------------ */
/* reset: terrain generation, joint noise, yaw, ground scan, FK          */

__global__ __launch_bounds__(64) void k_reset(DevWorlds d, const Model *__restrict__ m, const ResetK *__restrict__ rk,
                                              const int *__restrict__ ids, const uint32_t *__restrict__ seeds,
                                              const double *__restrict__ diff_d, const float *__restrict__ diff_f,
                                              int n_ids, const int *__restrict__ count_dev)
{
    __shared__ float hbuf[GRID_LIMIT_CELLS];
    __shared__ int8_t qbuf[(64 + 2 * TPAD) * (64 + 2 * TPAD)];
    __shared__ float lat[LAT_LIMIT];
    __shared__ float draws[NJ + 1];
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

    /* the draws are one sequential stream: lane 0 produces them */
    if (tid == 0)
        reset_draws(seeds[k], gen, n, lat, hbuf, draws);
    __syncthreads();
    /* quantisation steps of the columns (csrc/arc_terrain.c:54-71) */
    for (int c = tid; c < cells; c += 64)
        hbuf[c] = gen ? column_step(lat, n, dfc, hbuf[c], c) : 0.0f;
    __syncthreads();
    /* base: the pre-shift height of the column at the origin */
    float base = 0.0f;
    if (gen) {
        int ix = (int)floorf((0.0f - M.terr.ox) / M.terr.cell), iy = (int)floorf((0.0f - M.terr.oy) / M.terr.cell);
        base = 0.03f * hbuf[ix + iy * n];
    }
    int8_t *gq = d.tq + (size_t)i * d.tcells;
    for (int c = tid; c < d.tcells; c += 64) {
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

    if (tid == 0) {
        float ang[NJ];
        for (int j = 0; j < NJ; ++j)
            ang[j] = clampf_(R.def[j] + draws[j], R.lo[j], R.hi[j]);
        float ground = ground_under(M.terr, tl, R.footprint, R.cell);
        V3 bx[NB];
        Q4 bq[NB];
        pose_bodies(M, v3(0.0f, 0.0f, ground + R.stand + 0.03f), qaxis(v3(0.0f, 0.0f, 1.0f), draws[NJ]), ang, bx, bq);
        for (int b = 0; b < NB; ++b) {
            d.x[(b * 3 + 0) * S + i] = bx[b].x; d.x[(b * 3 + 1) * S + i] = bx[b].y; d.x[(b * 3 + 2) * S + i] = bx[b].z;
            d.q[(b * 4 + 0) * S + i] = bq[b].w; d.q[(b * 4 + 1) * S + i] = bq[b].x;
            d.q[(b * 4 + 2) * S + i] = bq[b].y; d.q[(b * 4 + 3) * S + i] = bq[b].z;
            for (int c = 0; c < 3; ++c) {
                d.v[(b * 3 + c) * S + i] = 0.0f;
                d.w[(b * 3 + c) * S + i] = 0.0f;
            }
        }
        for (int j = 0; j < NJ; ++j) {
            d.target[j * S + i] = ang[j];
            d.jout[(0 * NJ + j) * S + i] = ang[j];
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
    hipLaunchKernelGGL(k_reset, dim3(blocks), dim3(64), 0, s, d, m_dev, r_dev, ids, seeds, difficulty_d, difficulty_f,
                       n_ids, count_dev);
    return hipGetLastError();
}

/* fall rules of the task after a step, arc/task.py:180-203; comparisons in
 * binary64 on the binary32 state, as the task makes them on float64 copies */
__global__ void k_falls(DevWorlds d, const Model *__restrict__ m, TaskK tk, TaskState ts)
{
    const size_t S = d.stride;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.n)
        return;
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

hipError_t launch_falls(const DevWorlds &d, const Model *m_dev, const TaskK &tk, const TaskState &ts, hipStream_t s)
{
    if (d.n <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_falls, dim3((d.n + 255) / 256), dim3(256), 0, s, d, m_dev, tk, ts);
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

hipError_t launch_dump(const DevWorlds &d, int i, float *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_dump, dim3(1), dim3(64), 0, s, d, i, out);
    return hipGetLastError();
}

} /* namespace xpbd */
