/*
 * step_kernel.h -- the register-resident step kernels (one world per lane,
 * four or two lanes per world, four lanes with shared root solves; all
 * substeps of a control step in one launch) and their launcher, instantiated
 * once per workgroup size by step_b64.hip / step_b128.hip / step_b256.hip so
 * the three compile in parallel.
 *
 * Work distribution, per launch:
 *   plain      a grid-stride loop over worlds 0..total-1, targets world-major;
 *   batched    the same over a StepBatch read from device memory at run time
 *              (uneven world ranges of several requests, each range's targets
 *              read from the buffer its row pointer names; one captured graph
 *              serves any batch the host writes into that table);
 *   persistent a grid sized to the resident wavefronts, each wavefront taking
 *              the next 64 worlds from an atomic counter until none are left.
 * LDS layout, a template parameter: substep-start poses only (TILE = 0), or
 * poses plus a TILE x TILE terrain tile per world.
 */
#pragma once
#include "device_state.h"
#include "leg_engine.h"

namespace xpbd {

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
template <int BLOCK, int TILE> struct LdsTileCells {
    const TerrainK &t;
    const TerrainRef &r;
    const int8_t *lt;
    int tx, ty;
    __device__ __forceinline__ float operator()(int ix, int iy) const
    {
        unsigned u = (unsigned)(ix - tx), v = (unsigned)(iy - ty);
        if (ARC_LIKELY(u < (unsigned)TILE && v < (unsigned)TILE))
            return dequant(r, lt[(u + v * TILE) * BLOCK]);
        return terrain_cell(t, r, ix, iy);
    }
    /* a window of at most 2 x 2 columns inside the tile: true if no column's
     * top is above lo (the columns right of and above a one-column window are
     * read but not used); false when the window is not of that kind */
    __device__ __forceinline__ bool clear(int ix0, int ix1, int iy0, int iy1, float lo) const
    {
        const unsigned u = (unsigned)(ix0 - tx), v = (unsigned)(iy0 - ty);
        const bool wide = ix1 != ix0, tall = iy1 != iy0;
        if (ARC_UNLIKELY(!(u < (unsigned)(TILE - 1) && v < (unsigned)(TILE - 1) && ix1 - ix0 <= 1 && iy1 - iy0 <= 1)))
            return false;
        const int8_t *p = lt + (u + v * TILE) * BLOCK;
        const float h00 = dequant(r, p[0]), h10 = dequant(r, p[BLOCK]);
        const float h01 = dequant(r, p[TILE * BLOCK]), h11 = dequant(r, p[(TILE + 1) * BLOCK]);
        return lo >= h00 && (!wide || lo >= h10) && (!tall || lo >= h01) && (!(wide && tall) || lo >= h11);
    }
};
template <int BLOCK, int TILE> struct LdsTile {
    const int8_t *lt;
    int tx, ty;
    __device__ __forceinline__ LdsTileCells<BLOCK, TILE> cells(const TerrainK &t, const TerrainRef &r) const
    {
        return LdsTileCells<BLOCK, TILE>{ t, r, lt, tx, ty };
    }
};

/* fill the tile of world i around the torso position (x, y); lt = the
 * world's column 0 in LDS, columns STRIDE bytes apart */
template <int STRIDE, int TILE>
__device__ __forceinline__ LdsTile<STRIDE, TILE> fill_tile(int8_t *lt, const DevWorlds &d, const TerrainK &t,
                                                           size_t i, float x, float y)
{
    LdsTile<STRIDE, TILE> T;
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
                lt[(4 * w + b + v * TILE) * STRIDE] = (int8_t)(word >> (8 * b));
        }
    }
    return T;
}

/* Workgroups are handed to the 8 XCDs in turn (workgroup b to XCD b mod 8),
 * each XCD with its own L2.  With a grid that is a multiple of 8, workgroup b
 * takes the slot (b mod 8) (G / 8) + b / 8 of the grid instead of b, so that
 * each XCD steps one contiguous run of worlds and the cache lines of their
 * structure-of-arrays state are not split between two L2s (a quarter less
 * HBM fetched by the leg kernel at 65,536 worlds).  Any grid: a bijection. */
__device__ __forceinline__ int xcd_slot()
{
    const int G = (int)gridDim.x, b = (int)blockIdx.x;
    return G % 8 == 0 ? (b % 8) * (G / 8) + b / 8 : b;
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

/* one control step of world i by the calling lane */
template <int BLOCK, int TILE>
__device__ __forceinline__ void step_world(const DevWorlds &d, const Model &M, const float *__restrict__ row,
                                           LdsPoses<BLOCK> &P0, int8_t *s_tile, int i)
{
    const size_t S = d.stride;
    LaneWorld W;
    /* targets (row: the world's 12) clamped to the limits, csrc/arc_sim.c:132-134 */
    static_for<0, NJ>([&](auto j) {
        float tg = row[j];
        W.target[j] = clampf_(tg, M.joint[j].lo, M.joint[j].hi);
        d.target[j * S + i] = W.target[j];
    });
    if (d.unstable[i])
        return; /* frozen until reset, csrc/arc_sim.c:135-136 */
    GlobalAnchors A{ d.anc, d.ancn, S, (size_t)i };
    TerrainRef tr = terrain_ref(d, i);
    tr.base = d.tbase[i];
    tr.outside = d.outside[i];
    ContactStore cs;
    cs.base = d.cbuf;
    cs.stride = S;
    cs.world = i;
    cs.counts = d.ncontact;
    cs.cstride = S;
    cs.cworld = i;
    int ov;
    if constexpr (TILE > 0) {
        const LdsTile<BLOCK, TILE> tile =
            fill_tile<BLOCK, TILE>(s_tile + threadIdx.x, d, M.terr, (size_t)i, d.x[0 * S + i], d.x[1 * S + i]);
        load_world(d, i, W);
        ov = lane_step(W, A, P0, M, tr, cs, tile);
    } else {
        load_world(d, i, W);
        ov = lane_step(W, A, P0, M, tr, cs);
    }
    const size_t io = opaque_index((size_t)i);
    store_world(d, io, W);
    if (ov)
        d.err[io] |= ov;
}

template <int BLOCK, int TILE>
__global__ __launch_bounds__(BLOCK) void k_step(DevWorlds d, const Model *__restrict__ m,
                                                const float *__restrict__ targets, const StepBatch *__restrict__ batch,
                                                int total, int *__restrict__ work)
{
    __shared__ float s_pose[NB * 7 * BLOCK];
    __shared__ int8_t s_tile[TILE > 0 ? TILE * TILE * BLOCK : 1];
    const Model &M = *m;
    LdsPoses<BLOCK> P0;
    P0.p = s_pose;
    P0.tid = threadIdx.x;
    const int n_ranges = batch ? batch->n_ranges : 0;
    if (batch)
        total = batch->total;
    const int lane = threadIdx.x & 63;
    /* each wavefront steps 64 consecutive work items at a time: the next
     * ones from the counter (persistent) or one grid further on */
    auto next = [&](int base) {
        if (!work)
            return base + (int)gridDim.x * BLOCK;
        int v = 0;
        if (lane == 0)
            v = atomicAdd(work, 64);
        return __shfl(v, 0, 64);
    };
    int base = work ? next(0) : xcd_slot() * BLOCK + (threadIdx.x & ~63);
    while (base < total) {
        const int t = base + lane;
        if (t < total) {
            const float *row = nullptr;
            const int i = batch ? batch_world(batch, n_ranges, t, row) : t;
            if (!batch)
                row = targets + (size_t)t * NJ;
            step_world<BLOCK, TILE>(d, M, row, P0, s_tile, i);
        }
        base = next(base);
    }
}

/* ------------------------------------------------------------------ */
/* group forms: the P = 4 / K lanes of a group step one world, K legs per
 * lane (leg_engine.h) */

template <int K>
__device__ __forceinline__ void load_leg(const DevWorlds &d, size_t i, int q, LegWorld<K> &W)
{
    const size_t S = d.stride;
    auto body = [&](int b, Body &B) {
        B.x = v3(d.x[(b * 3 + 0) * S + i], d.x[(b * 3 + 1) * S + i], d.x[(b * 3 + 2) * S + i]);
        B.q = q4(d.q[(b * 4 + 0) * S + i], d.q[(b * 4 + 1) * S + i], d.q[(b * 4 + 2) * S + i], d.q[(b * 4 + 3) * S + i]);
        B.v = v3(d.v[(b * 3 + 0) * S + i], d.v[(b * 3 + 1) * S + i], d.v[(b * 3 + 2) * S + i]);
        B.w = v3(d.w[(b * 3 + 0) * S + i], d.w[(b * 3 + 1) * S + i], d.w[(b * 3 + 2) * S + i]);
    };
    body(0, W.T);
    static_for<0, K>([&](auto l) {
        const int g = q * K + l;
        body(1 + 3 * g, W.b[l][0]);
        body(2 + 3 * g, W.b[l][1]);
        body(3 + 3 * g, W.b[l][2]);
    });
    const uint32_t a = d.anchored[i];
    constexpr uint32_t mask = (1u << (5 * K)) - 1u;
    W.anchored = (a & 1u) | (((a >> (5 * q * K + 1)) & mask) << 1);
    W.unstable = d.unstable[i];
    W.time = d.time[i];
}

template <int K>
__device__ __forceinline__ void store_leg(const DevWorlds &d, size_t i, int q, const LegWorld<K> &W, int ov)
{
    constexpr int P = LEGS / K;
    const size_t S = d.stride;
    auto body = [&](int b, const Body &B) {
        d.x[(b * 3 + 0) * S + i] = B.x.x; d.x[(b * 3 + 1) * S + i] = B.x.y; d.x[(b * 3 + 2) * S + i] = B.x.z;
        d.q[(b * 4 + 0) * S + i] = B.q.w; d.q[(b * 4 + 1) * S + i] = B.q.x;
        d.q[(b * 4 + 2) * S + i] = B.q.y; d.q[(b * 4 + 3) * S + i] = B.q.z;
        d.v[(b * 3 + 0) * S + i] = B.v.x; d.v[(b * 3 + 1) * S + i] = B.v.y; d.v[(b * 3 + 2) * S + i] = B.v.z;
        d.w[(b * 3 + 0) * S + i] = B.w.x; d.w[(b * 3 + 1) * S + i] = B.w.y; d.w[(b * 3 + 2) * S + i] = B.w.z;
    };
    static_for<0, K>([&](auto l) {
        const int g = q * K + l;
        body(1 + 3 * g, W.b[l][0]);
        body(2 + 3 * g, W.b[l][1]);
        body(3 + 3 * g, W.b[l][2]);
        for (int m = 0; m < 3; ++m) {
            const int j = 3 * g + m;
            d.jout[(0 * NJ + j) * S + i] = W.angle[3 * l + m];
            d.jout[(1 * NJ + j) * S + i] = W.rate[3 * l + m];
            d.jout[(2 * NJ + j) * S + i] = W.torque[3 * l + m];
        }
        for (int m = 0; m < 5; ++m)
            d.touch[(5 * g + 1 + m) * S + i] = (uint8_t)leg_touch(W, 1 + 5 * l + m);
    });
    constexpr uint32_t mask = (1u << (5 * K)) - 1u;
    const int anch = gor<P>((int)(((W.anchored >> 1) & mask) << (5 * q * K + 1)));
    const int err = gor<P>(ov);
    if (q == 0) {
        body(0, W.T);
        d.touch[i] = (uint8_t)leg_touch(W, 0);
        d.anchored[i] = (uint32_t)anch | (W.anchored & 1u);
        d.unstable[i] = W.unstable;
        d.time[i] = W.time;
        if (err)
            d.err[i] |= err;
    }
}

/* the root joints of the group forms: a lane pair per root with one leg per
 * lane, the owner lane with two */
template <int K> using LegRoots = std::conditional_t<K == 1, RootsPair, RootsLocal<K>>;

/* one control step of world i by lane q of its group */
template <int BLOCK, int TILE, int K>
__device__ __forceinline__ void step_world_leg(const DevWorlds &d, const Model &M, const float *__restrict__ row,
                                               LdsPoses<BLOCK> &P0, int8_t *s_tile, int i, int q, float *crec,
                                               const float *s_jgeo)
{
    constexpr int P = LEGS / K;
    const size_t S = d.stride;
    LegWorld<K> W;
    for (int m = 0; m < 3 * K; ++m) {
        const int j = 3 * q * K + m;
        float tg = row[j];
        W.target[m] = clampf_(tg, M.joint[j].lo, M.joint[j].hi);
        d.target[j * S + i] = W.target[m];
    }
    if (d.unstable[i])
        return; /* frozen until reset, csrc/arc_sim.c:135-136 */
    const size_t So = own_sgpr(S);
    GlobalAnchors A{ own_sgpr(d.anc), own_sgpr(d.ancn), So, (size_t)i };
    TerrainRef tr = terrain_ref(d, i);
    tr.base = d.tbase[i];
    tr.outside = d.outside[i];
    ContactStore cs;
    cs.base = own_sgpr(d.lcbuf);
    cs.stride = P * So;
    cs.world = (size_t)i * P + q;
    cs.mirror = crec;
    cs.mstride = BLOCK;
    cs.counts = own_sgpr(d.ncontact);
    cs.cstride = So;
    cs.cworld = (size_t)i;
    constexpr int WPB = BLOCK / P;
    int ov;
    if constexpr (TILE > 0) {
        /* the group's lanes write the same bytes: no ordering needed */
        const LdsTile<WPB, TILE> tile = fill_tile<WPB, TILE>(s_tile + threadIdx.x / P, d, M.terr, (size_t)i,
                                                             d.x[0 * S + i], d.x[1 * S + i]);
        load_leg<K>(d, i, q, W);
        LegRoots<K> roots;
        ov = leg_step<K>(W, q, A, P0, M, tr, cs, tile, roots, LdsJoints{ s_jgeo });
    } else {
        load_leg<K>(d, i, q, W);
        LegRoots<K> roots;
        ov = leg_step<K>(W, q, A, P0, M, tr, cs, GridTile(), roots, LdsJoints{ s_jgeo });
    }
    store_leg<K>(d, opaque_index((size_t)i), q, W, ov);
}

template <int BLOCK, int TILE, int K>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(K == 1 ? 2 : 1))) void k_step_leg(
    DevWorlds d, const Model *__restrict__ m, const float *__restrict__ targets, const StepBatch *__restrict__ batch,
    int total, int *__restrict__ work)
{
    constexpr int P = LEGS / K;
    __shared__ float s_pose[(1 + 3 * K) * 7 * BLOCK];
    __shared__ int8_t s_tile[TILE > 0 ? TILE * TILE * (BLOCK / P) : 1];
    __shared__ float s_crec[K == 1 ? CMIRROR * 6 * BLOCK : 1]; /* each lane's first tibia contact records */
    __shared__ float s_jgeo[2 * LEGS * JGEO];                  /* the legs' mid and distal joint geometry */
    fill_joint_table(s_jgeo);
    __syncthreads();
    const Model &M = *m;
    LdsPoses<BLOCK> P0;
    P0.p = s_pose;
    P0.tid = threadIdx.x;
    const int n_ranges = batch ? batch->n_ranges : 0;
    if (batch)
        total = batch->total;
    const int items = P * total;
    const int lane = threadIdx.x & 63;
    auto next = [&](int base) {
        if (!work)
            return base + (int)gridDim.x * BLOCK;
        int v = 0;
        if (lane == 0)
            v = atomicAdd(work, 64);
        return __shfl(v, 0, 64);
    };
    int base = work ? next(0) : xcd_slot() * BLOCK + (threadIdx.x & ~63);
    while (base < items) {
        const int t = base + lane;
        if (t < items) {
            const int w = t / P;
            const float *row = nullptr;
            const int i = batch ? batch_world(batch, n_ranges, w, row) : w;
            if (!batch)
                row = targets + (size_t)w * NJ;
            step_world_leg<BLOCK, TILE, K>(d, M, row, P0, s_tile, i, t % P, K == 1 ? s_crec + threadIdx.x : nullptr,
                                           s_jgeo);
        }
        base = next(base);
    }
}

/* Four lanes per world with the root joints of the block's BLOCK / 4 worlds
 * solved by the first lanes of one wavefront (RootsShared).  Every lane of
 * the block takes part in the exchanges, so lanes without a world of their own
 * (the last block's tail, worlds frozen by instability) step a copy of a real
 * world's state with their anchors and contact lists in the dummy area and
 * store nothing. */
template <int BLOCK>
__device__ __forceinline__ void step_world_shared(const DevWorlds &d, const Model &M, const float *__restrict__ row,
                                                  LdsPoses<BLOCK> &P0, int8_t *s_tile, RootsShared<BLOCK / 4> &roots,
                                                  int i, int q, bool valid)
{
    constexpr int WPB = BLOCK / 4;
    const size_t S = d.stride;
    LegWorld<1> W;
    for (int m = 0; m < 3; ++m) {
        const int j = 3 * q + m;
        W.target[m] = clampf_(row[j], M.joint[j].lo, M.joint[j].hi);
        if (valid)
            d.target[j * S + i] = W.target[m];
    }
    const bool live = valid && !d.unstable[i];
    /* the dummy area holds 256 lanes' lists and anchors */
    const float *dummy_anc = d.lcdummy + (size_t)(M.ccap[0] + M.ccap[1] + M.ccap[5]) * 6 * 256;
    GlobalAnchors A = live ? GlobalAnchors{ d.anc, d.ancn, S, (size_t)i }
                           : GlobalAnchors{ (float *)dummy_anc, (float *)dummy_anc + NS * 3 * 256, (size_t)256,
                                            (size_t)threadIdx.x };
    TerrainRef tr = terrain_ref(d, i);
    tr.base = d.tbase[i];
    tr.outside = d.outside[i];
    ContactStore cs;
    cs.base = live ? d.lcbuf : d.lcdummy;
    cs.stride = live ? 4 * S : (size_t)256;
    cs.world = live ? (size_t)i * 4 + q : (size_t)threadIdx.x;
    cs.counts = live ? d.ncontact : nullptr; /* lanes without a world of their own store nothing */
    cs.cstride = S;
    cs.cworld = (size_t)i;
    const LdsTile<WPB, 16> tile = fill_tile<WPB, 16>(s_tile + threadIdx.x / 4, d, M.terr, (size_t)i, d.x[0 * S + i],
                                                     d.x[1 * S + i]);
    load_leg<1>(d, i, q, W);
    const int ov = leg_step<1>(W, q, A, P0, M, tr, cs, tile, roots);
    if (live)
        store_leg<1>(d, opaque_index((size_t)i), q, W, ov);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(2))) void k_step_shared(
    DevWorlds d, const Model *__restrict__ m, const float *__restrict__ targets, const StepBatch *__restrict__ batch,
    int total)
{
    static_assert(BLOCK == 128 || BLOCK == 256, "blocks of two or four wavefronts");
    constexpr int WPB = BLOCK / 4;
    __shared__ float s_pose[4 * 7 * BLOCK];
    __shared__ int8_t s_tile[16 * 16 * WPB];
    __shared__ float s_roots[RS_FIELDS * WPB];
    __shared__ float s_rootv[RS_VFIELDS * WPB];
    const Model &M = *m;
    LdsPoses<BLOCK> P0;
    P0.p = s_pose;
    P0.tid = threadIdx.x;
    RootsShared<WPB> roots;
    roots.xs = s_roots;
    roots.xv = s_rootv;
    roots.g = threadIdx.x >> 2;
    const int n_ranges = batch ? batch->n_ranges : 0;
    if (batch)
        total = batch->total;
    if (total <= 0)
        return;
    const int items = 4 * total;
    /* whole blocks move together: every lane makes the same exchanges */
    for (int base = xcd_slot() * BLOCK; base < items; base += (int)gridDim.x * BLOCK) {
        const int t = base + (int)threadIdx.x;
        const bool valid = t < items;
        const int w = (valid ? t : items - 1) >> 2;
        const float *row = nullptr;
        const int i = batch ? batch_world(batch, n_ranges, w, row) : w;
        if (!batch)
            row = targets + (size_t)w * NJ;
        step_world_shared<BLOCK>(d, M, row, P0, s_tile, roots, i, (int)threadIdx.x & 3, valid);
    }
}

/* workgroups of the BLOCK-thread kernel resident on the whole device at once */
template <int BLOCK>
int step_resident_blocks_of(int tile, int leg)
{
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess
        || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    hipError_t e;
    if (leg == 4) {
        if constexpr (BLOCK == 128 || BLOCK == 256)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_step_shared<BLOCK>, BLOCK, 0);
        else
            e = hipErrorInvalidValue;
    }
    else if (leg == 1)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_step_leg<BLOCK, 16, 1>, BLOCK, 0);
    else if (leg == 2)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_step_leg<BLOCK, 16, 2>, BLOCK, 0);
    else
        e = tile ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_step<BLOCK, 16>, BLOCK, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_step<BLOCK, 0>, BLOCK, 0);
    if (e != hipSuccess)
        return 0;
    return per * cus;
}

/* launcher of the BLOCK-thread kernels (grid_blocks 0: one item per thread) */
template <int BLOCK>
hipError_t launch_step_block(const DevWorlds &d, const Model *m_dev, const float *targets, const StepBatch *batch,
                             int total, int tile, int grid_blocks, int *work, int leg, hipStream_t s)
{
    const int per_world = leg == 1 || leg == 4 ? 4 : leg == 2 ? 2 : 1;
    if (grid_blocks <= 0)
        grid_blocks = (per_world * total + BLOCK - 1) / BLOCK;
    if (grid_blocks <= 0)
        return hipSuccess;
    if (leg) {
        /* the group forms always hold the terrain tile (the autotuner never
         * picked them without it) */
        if (!d.lcbuf || tile != 16)
            return hipErrorInvalidValue;
        if (leg == 4) {
            /* whole blocks only; no work counter (the exchanges need every lane) */
            if constexpr (BLOCK == 128 || BLOCK == 256) {
                if (work || !d.lcdummy)
                    return hipErrorInvalidValue;
                hipLaunchKernelGGL(k_step_shared<BLOCK>, dim3(grid_blocks), dim3(BLOCK), 0, s, d, m_dev, targets,
                                   batch, total);
                return hipGetLastError();
            } else {
                return hipErrorInvalidValue;
            }
        }
        if (leg == 1)
            hipLaunchKernelGGL((k_step_leg<BLOCK, 16, 1>), dim3(grid_blocks), dim3(BLOCK), 0, s, d, m_dev, targets,
                               batch, total, work);
        else if (leg == 2)
            hipLaunchKernelGGL((k_step_leg<BLOCK, 16, 2>), dim3(grid_blocks), dim3(BLOCK), 0, s, d, m_dev, targets,
                               batch, total, work);
        else
            return hipErrorInvalidValue;
        return hipGetLastError();
    }
    if (tile == 16)
        hipLaunchKernelGGL((k_step<BLOCK, 16>), dim3(grid_blocks), dim3(BLOCK), 0, s, d, m_dev, targets, batch, total,
                           work);
    else if (tile == 0)
        hipLaunchKernelGGL((k_step<BLOCK, 0>), dim3(grid_blocks), dim3(BLOCK), 0, s, d, m_dev, targets, batch, total,
                           work);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

} /* namespace xpbd */
