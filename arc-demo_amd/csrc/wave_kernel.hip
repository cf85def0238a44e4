/*
 * wave_kernel.hip -- the plain form of the step: one world per wavefront,
 * one launch per control step.
 *
 * The world's bodies, joint accumulators, sphere anchors and the substep's
 * contact list are staged in LDS.  Lanes work in parallel where the reference
 * loop iterations are independent:
 *   - integration and velocity update: one lane per body;
 *   - joint projection (csrc/arc_phys.c:729-730, 734-735): Gauss-Seidel in the
 *     reference order, executed level by level of its dependency DAG -- a
 *     joint only reads and writes its parent and child, so joints with no body
 *     in common commute and each level runs on parallel lanes, giving the
 *     serial order's results bit for bit;
 *   - sphere contacts: one lane per body, each walking that body's spheres in
 *     order (a contact solve moves only the sphere's own body,
 *     csrc/arc_phys.c:488-500); where the spheres' cell windows could exceed
 *     the reference's 128-record cap, one lane walks every sphere in the
 *     world's order into a single list, so the cap cuts where it does there;
 *   - contact velocities: one lane per body, its contacts in record order.
 * The topology comes from tables built on the host from the world itself, and
 * every body is the run-time-kinded KGEN, so the code follows the reference's
 * generic arithmetic literally.
 */
#include "device_state.h"
#include "wave_tables.h"
#include "world_step.h"

namespace xpbd {

struct WaveWorld {
    Body b[WAVE_MAX_BODIES];
    float target[WAVE_MAX_JOINTS], tacc[WAVE_MAX_JOINTS];
    float angle[WAVE_MAX_JOINTS], rate[WAVE_MAX_JOINTS], torque[WAVE_MAX_JOINTS];
    int touch[WAVE_MAX_SPHERES], anchored[WAVE_MAX_SPHERES];
    V3 anc[WAVE_MAX_SPHERES], ancn[WAVE_MAX_SPHERES];
    Pose0 p0[WAVE_MAX_BODIES];
    float cl[WAVE_MAX_SLOTS][6];
    int cbody[WAVE_MAX_SLOTS]; /* serial collision: body of each record of the one list */
    int ccount[WAVE_MAX_BODIES];
    int ntotal;                /* serial collision: records in the list */
    int unstable;
};

template <int PK, class MT>
__device__ __forceinline__ void wave_joint_pos(WaveWorld &W, const MT &M, const WaveK &T, int j)
{
    int p = T.jp[j], c = T.jc[j];
    Body C = W.b[c];
    if (PK >= 0) {
        Body P = W.b[p];
        joint_positions<KGEN, KGEN>(&P, &M.body[p], C, M.body[c], M.joint[j], W.target[j], M.hh, W.tacc[j]);
        W.b[p] = P;
    } else {
        joint_positions<-1, KGEN>(nullptr, nullptr, C, M.body[c], M.joint[j], W.target[j], M.hh, W.tacc[j]);
    }
    W.b[c] = C;
}

template <int PK, class MT>
__device__ __forceinline__ void wave_joint_vel(WaveWorld &W, const MT &M, const WaveK &T, int j)
{
    int p = T.jp[j], c = T.jc[j];
    Body C = W.b[c];
    if (PK >= 0) {
        Body P = W.b[p];
        joint_velocities<KGEN, KGEN>(&P, &M.body[p], C, M.body[c], M.joint[j], M.h, W.tacc[j]);
        W.b[p] = P;
    } else {
        joint_velocities<-1, KGEN>(nullptr, nullptr, C, M.body[c], M.joint[j], M.h, W.tacc[j]);
    }
    W.b[c] = C;
}

/* spheres sph[0 .. count) of contact-body group g, in that order, against the
 * terrain whose columns `cells` returns.  Records go to the group's own slot
 * range, or, with `serial`, to the world's one list in record order (at most
 * MAXC records, the reference's cap) tagged with their body. */
template <class MT, class Cells>
__device__ __forceinline__ void wave_collide(WaveWorld &W, const MT &M, const WaveK &T, int g, const int *sph, int count,
                                             const Cells &cells, int &total, bool serial)
{
    const int B = T.cb_body[g];
    Body b = W.b[B];
    int &cnt = W.ccount[B];
    for (int k = 0; k < count; ++k) {
        const int s = sph[k];
        auto rec = [&](V3 l, V3 n) {
            if (total < MAXC) {
                float *c = serial ? W.cl[total] : cnt < T.cb_cap[g] ? W.cl[T.cb_slot[g] + cnt] : nullptr;
                if (c) {
                    c[0] = l.x; c[1] = l.y; c[2] = l.z; c[3] = n.x; c[4] = n.y; c[5] = n.z;
                }
                if (serial)
                    W.cbody[total] = B;
                cnt++;
                total++;
            }
        };
        AnchorRef an{ W.anc[s], W.ancn[s] };
        auto pose0 = [&]() { return W.p0[B]; };
        collide<KGEN, AT_GEN>(b, M.body[B], M.sphere[s].at, M.sphere[s].rad, W.touch[s], W.anchored[s], an, pose0,
                              M.terr, cells, rec, M.mus, M.mud);
    }
    W.b[B] = b;
}

/* The substeps of one world step (csrc/arc_phys.c:727-737) on the world
 * staged in W, by the 64 lanes of the calling wavefront.  The caller has run
 * the step's prologue (caches, touch counts and torque sums cleared). */
template <class MT, class Cells>
__device__ __forceinline__ void wave_substeps(WaveWorld &W, const MT &M, const WaveK &T, const Cells &cells, int lane)
{
    for (int sub = 0; sub < M.substeps; ++sub) {
        if (lane < T.nb) {
            Body b = W.b[lane];
            W.p0[lane] = integrate<KGEN>(b, M.body[lane], M.g, M.h);
            W.b[lane] = b;
            W.ccount[lane] = 0;
        }
        __syncthreads();
        for (int L = 0; L < T.nlev; ++L) {
            int k = T.lev_first[L] + lane;
            if (k < T.lev_first[L + 1]) {
                int j = T.lev_joint[k];
                if (T.jp[j] >= 0)
                    wave_joint_pos<KGEN>(W, M, T, j);
                else
                    wave_joint_pos<-1>(W, M, T, j);
            }
            __syncthreads();
        }
        if (T.serial_collide) {
            /* the contact-record cap could bind: every sphere in the world's
             * sphere order, as the reference records them */
            if (lane == 0) {
                int total = 0;
                for (int s = 0; s < T.ns; ++s)
                    wave_collide(W, M, T, T.s_group[s], &s, 1, cells, total, true);
                W.ntotal = total;
            }
        } else if (lane < T.ncb) {
            int total = 0; /* the cap cannot bind: groups are independent */
            wave_collide(W, M, T, lane, &T.cb_sphere[T.cb_first[lane]], T.cb_first[lane + 1] - T.cb_first[lane], cells,
                         total, false);
        }
        __syncthreads();
        if (lane < T.nb) {
            Body b = W.b[lane];
            velocities(b, W.p0[lane], M.inv_h, M.two_h);
            W.b[lane] = b;
        }
        __syncthreads();
        for (int L = 0; L < T.nlev; ++L) {
            int k = T.lev_first[L] + lane;
            if (k < T.lev_first[L + 1]) {
                int j = T.lev_joint[k];
                if (T.jp[j] >= 0)
                    wave_joint_vel<KGEN>(W, M, T, j);
                else
                    wave_joint_vel<-1>(W, M, T, j);
            }
            __syncthreads();
        }
        if (lane < T.ncb) {
            const int B = T.cb_body[lane];
            Body b = W.b[B];
            if (T.serial_collide) {
                /* the body's records in record order */
                for (int k = 0; k < W.ntotal; ++k)
                    if (W.cbody[k] == B) {
                        const float *c = W.cl[k];
                        contact_velocity<KGEN>(b, M.body[B], v3(c[0], c[1], c[2]), v3(c[3], c[4], c[5]));
                    }
            } else {
                int n = W.ccount[B] < T.cb_cap[lane] ? W.ccount[B] : T.cb_cap[lane];
                for (int k = 0; k < n; ++k) {
                    const float *c = W.cl[T.cb_slot[lane] + k];
                    contact_velocity<KGEN>(b, M.body[B], v3(c[0], c[1], c[2]), v3(c[3], c[4], c[5]));
                }
            }
            W.b[B] = b;
        }
        __syncthreads();
    }
}

/* joint outputs of the step's epilogue (csrc/arc_phys.c:739-744) */
template <class MT>
__device__ __forceinline__ void wave_joint_outputs(WaveWorld &W, const MT &M, const WaveK &T, int lane)
{
    if (lane < T.nj) {
        int p = T.jp[lane], c = T.jc[lane];
        Body C = W.b[c];
        if (p >= 0) {
            Body P = W.b[p];
            W.angle[lane] = joint_angle<true>(&P, C, M.joint[lane]);
            W.rate[lane] = joint_rate<true>(&P, C, M.joint[lane]);
        } else {
            W.angle[lane] = joint_angle<false>(nullptr, C, M.joint[lane]);
            W.rate[lane] = joint_rate<false>(nullptr, C, M.joint[lane]);
        }
        W.torque[lane] = W.tacc[lane] / (float)M.substeps;
    }
}

template <class TT>
__global__ __launch_bounds__(64) void k_step_wave(DevWorlds d, const Model *__restrict__ m, const WaveK *__restrict__ wk,
                                                  const TT *__restrict__ targets, int total_worlds)
{
    __shared__ WaveWorld W;
    const Model &M = *m;
    const WaveK &T = *wk;
    const int lane = threadIdx.x;
    const size_t S = d.stride;
    for (int i = blockIdx.x; i < total_worlds; i += gridDim.x) {
        /* targets clamped to the limits (every world), csrc/arc_sim.c:132-134 */
        if (lane < T.nj) {
            float tg = clampf_((float)targets[(size_t)i * T.nj + lane], M.joint[lane].lo, M.joint[lane].hi);
            W.target[lane] = tg;
            d.target[lane * S + i] = tg;
        }
        if (lane == 0)
            W.unstable = d.unstable[i];
        __syncthreads();
        if (W.unstable) {
            __syncthreads();
            continue;
        }
        /* stage the world in LDS; prologue of arc_world_step */
        if (lane < T.nb) {
            Body &b = W.b[lane];
            b.x = v3(d.x[(lane * 3 + 0) * S + i], d.x[(lane * 3 + 1) * S + i], d.x[(lane * 3 + 2) * S + i]);
            b.q = q4(d.q[(lane * 4 + 0) * S + i], d.q[(lane * 4 + 1) * S + i], d.q[(lane * 4 + 2) * S + i],
                     d.q[(lane * 4 + 3) * S + i]);
            b.v = v3(d.v[(lane * 3 + 0) * S + i], d.v[(lane * 3 + 1) * S + i], d.v[(lane * 3 + 2) * S + i]);
            b.w = v3(d.w[(lane * 3 + 0) * S + i], d.w[(lane * 3 + 1) * S + i], d.w[(lane * 3 + 2) * S + i]);
            Body t = b;
            bsync<KGEN>(t, M.body[lane]);
            b = t;
            W.ccount[lane] = 0;
        }
        if (lane < T.ns) {
            uint32_t an = d.anchored[i];
            W.touch[lane] = 0;
            W.anchored[lane] = (an >> lane) & 1u;
            W.anc[lane] = v3(d.anc[(lane * 3 + 0) * S + i], d.anc[(lane * 3 + 1) * S + i], d.anc[(lane * 3 + 2) * S + i]);
            W.ancn[lane] = v3(d.ancn[(lane * 3 + 0) * S + i], d.ancn[(lane * 3 + 1) * S + i], d.ancn[(lane * 3 + 2) * S + i]);
        }
        if (lane < T.nj)
            W.tacc[lane] = 0.0f;
        __syncthreads();

        TerrainRef tr = terrain_ref(d, i);
        tr.base = d.tbase[i];
        tr.outside = d.outside[i];

        wave_substeps(W, M, T, GridCells{ M.terr, tr }, lane);
        wave_joint_outputs(W, M, T, lane);
        bool bad = lane < T.nb && body_unstable(W.b[lane]);
        bool any_bad = __any(bad);
        __syncthreads();
        /* write back */
        if (lane < T.nb) {
            const Body &b = W.b[lane];
            d.x[(lane * 3 + 0) * S + i] = b.x.x; d.x[(lane * 3 + 1) * S + i] = b.x.y; d.x[(lane * 3 + 2) * S + i] = b.x.z;
            d.q[(lane * 4 + 0) * S + i] = b.q.w; d.q[(lane * 4 + 1) * S + i] = b.q.x;
            d.q[(lane * 4 + 2) * S + i] = b.q.y; d.q[(lane * 4 + 3) * S + i] = b.q.z;
            d.v[(lane * 3 + 0) * S + i] = b.v.x; d.v[(lane * 3 + 1) * S + i] = b.v.y; d.v[(lane * 3 + 2) * S + i] = b.v.z;
            d.w[(lane * 3 + 0) * S + i] = b.w.x; d.w[(lane * 3 + 1) * S + i] = b.w.y; d.w[(lane * 3 + 2) * S + i] = b.w.z;
        }
        if (lane < T.nj) {
            d.jout[(0 * NJ + lane) * S + i] = W.angle[lane];
            d.jout[(1 * NJ + lane) * S + i] = W.rate[lane];
            d.jout[(2 * NJ + lane) * S + i] = W.torque[lane];
        }
        if (lane < T.ns) {
            d.touch[lane * S + i] = (uint8_t)W.touch[lane];
            d.anc[(lane * 3 + 0) * S + i] = W.anc[lane].x; d.anc[(lane * 3 + 1) * S + i] = W.anc[lane].y;
            d.anc[(lane * 3 + 2) * S + i] = W.anc[lane].z;
            d.ancn[(lane * 3 + 0) * S + i] = W.ancn[lane].x; d.ancn[(lane * 3 + 1) * S + i] = W.ancn[lane].y;
            d.ancn[(lane * 3 + 2) * S + i] = W.ancn[lane].z;
        }
        unsigned long long an = __ballot(lane < T.ns && W.anchored[lane]);
        int ov = 0;
        if (lane < T.ncb) {
            ov = W.ccount[T.cb_body[lane]] > T.cb_cap[lane] ? (1 << lane) : 0;
            /* the last substep's contacts, for the host snapshot: copied from
             * LDS into the one-world-per-lane form's list layout */
            const int B = T.cb_body[lane], ci = cbody_index(B);
            int n = 0;
            if (T.serial_collide) {
                for (int k = 0; k < W.ntotal && n < M.ccap[ci]; ++k)
                    if (W.cbody[k] == B) {
                        for (int f = 0; f < 6; ++f)
                            d.cbuf[((size_t)(M.cslot[ci] + n) * 6 + f) * S + i] = W.cl[k][f];
                        ++n;
                    }
            } else {
                n = W.ccount[B] < T.cb_cap[lane] ? W.ccount[B] : T.cb_cap[lane];
                for (int k = 0; k < n; ++k)
                    for (int f = 0; f < 6; ++f)
                        d.cbuf[((size_t)(M.cslot[ci] + k) * 6 + f) * S + i] = W.cl[T.cb_slot[lane] + k][f];
            }
            d.ncontact[(size_t)ci * S + i] = (uint8_t)n;
        }
        for (int off = 32; off > 0; off >>= 1)
            ov |= __shfl_xor(ov, off);
        if (lane == 0) {
            d.anchored[i] = (uint32_t)an;
            d.time[i] = d.time[i] + M.dt;
            if (any_bad)
                d.unstable[i] = 1;
            if (ov)
                d.err[i] |= ov;
        }
        __syncthreads();
    }
}

hipError_t launch_step_wave(const DevWorlds &d, const Model *m_dev, const WaveK *wk_dev, const double *targets_d,
                            const float *targets_f, int grid_blocks, hipStream_t s)
{
    if (d.n <= 0)
        return hipSuccess;
    if (grid_blocks <= 0)
        grid_blocks = d.n;
    if (targets_d)
        hipLaunchKernelGGL(k_step_wave<double>, dim3(grid_blocks), dim3(64), 0, s, d, m_dev, wk_dev, targets_d, d.n);
    else
        hipLaunchKernelGGL(k_step_wave<float>, dim3(grid_blocks), dim3(64), 0, s, d, m_dev, wk_dev, targets_f, d.n);
    return hipGetLastError();
}

/* columns of a terrain stored as binary32 heights, nx * ny row-major */
struct FloatCells {
    const TerrainK &t;
    const float *h;
    float outside;
    __device__ float operator()(int ix, int iy) const
    {
        if (ix < 0 || iy < 0 || ix >= t.nx || iy >= t.ny)
            return outside;
        return h[ix + iy * t.nx];
    }
};

/* `steps` steps of one hand-built world: the reference's arc_world_step
 * (csrc/arc_phys.c:707-747) on the world's own joint targets, unclamped, with
 * no skip of an unstable world; the state stays in LDS between steps */
__global__ __launch_bounds__(64) void k_world_step(WorldState *__restrict__ st, const WorldModel *__restrict__ m,
                                                   const WaveK *__restrict__ wk, const float *__restrict__ heights,
                                                   int steps)
{
    __shared__ WaveWorld W;
    const WorldModel &M = *m;
    const WaveK &T = *wk;
    const int lane = threadIdx.x;
    if (lane < T.nb) {
        Body &b = W.b[lane];
        const Body &s = st->b[lane];
        b.x = s.x;
        b.q = s.q;
        b.v = s.v;
        b.w = s.w;
    }
    if (lane < T.nj)
        W.target[lane] = st->target[lane];
    if (lane < T.ns) {
        W.touch[lane] = st->touch[lane];
        W.anchored[lane] = st->anchored[lane];
        W.anc[lane] = st->anc[lane];
        W.ancn[lane] = st->ancn[lane];
    }
    float time = st->time;
    int unstable = st->unstable, ov = 0;
    const FloatCells cells{ M.terr, heights, M.outside };
    for (int k = 0; k < steps; ++k) {
        /* prologue: the caller may have changed orientations since the last step */
        if (lane < T.nb) {
            Body t = W.b[lane];
            bsync<KGEN>(t, M.body[lane]);
            W.b[lane] = t;
        }
        if (lane < T.ns)
            W.touch[lane] = 0;
        if (lane < T.nj)
            W.tacc[lane] = 0.0f;
        __syncthreads();
        wave_substeps(W, M, T, cells, lane);
        wave_joint_outputs(W, M, T, lane);
        time = time + M.dt;
        if (__any(lane < T.nb && body_unstable(W.b[lane])))
            unstable = 1;
        if (__any(lane < T.ncb && W.ccount[T.cb_body[lane]] > T.cb_cap[lane]))
            ov = 1;
        __syncthreads();
    }
    if (lane < T.nb) {
        Body &s = st->b[lane];
        const Body &b = W.b[lane];
        s.x = b.x;
        s.q = b.q;
        s.v = b.v;
        s.w = b.w;
    }
    if (lane < T.nj) {
        st->angle[lane] = W.angle[lane];
        st->rate[lane] = W.rate[lane];
        st->torque[lane] = W.torque[lane];
    }
    if (lane < T.ns) {
        st->touch[lane] = W.touch[lane];
        st->anchored[lane] = W.anchored[lane];
        st->anc[lane] = W.anc[lane];
        st->ancn[lane] = W.ancn[lane];
    }
    if (lane == 0) {
        st->time = time;
        st->unstable = unstable;
        st->overflow |= ov;
    }
}

hipError_t launch_world_step(WorldState *st, const WorldModel *m, const WaveK *wk, const float *heights, int steps,
                             hipStream_t s)
{
    if (steps <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_world_step, dim3(1), dim3(64), 0, s, st, m, wk, heights, steps);
    return hipGetLastError();
}

} /* namespace xpbd */
