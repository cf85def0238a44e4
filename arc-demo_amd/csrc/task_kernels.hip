/*
 * task_kernels.hip -- the checkpoint-reaching task on the device, one thread
 * per world: actions to drive targets, rewards, terminations, checkpoints,
 * the terrain curriculum and the observation vector (arc/task.py:155-259),
 * with the vectorised env's auto-reset order (arc/env.py:99-117).
 *
 * The task's arithmetic is binary64 on the binary32 engine state, as the
 * reference task computes on float64 copies of the sim's outputs; its
 * transcendental functions are the device's binary64 ones.  Checkpoint
 * placement and reset seeds draw either from a per-world xorshift128 stream
 * of the engine's generator, or (numpy mode) from one numpy default_rng
 * stream in the reference's order: the hits of a step in world order, then
 * the seeds of the ended worlds, then their first checkpoints
 * (arc/task.py:112-150,232-234, arc/env.py:99-117).
 */
#include "task_state.h"

namespace xpbd {

/* binary64 view of the state observe() reports (csrc/arc_sim.c:140-183) */
struct TaskView {
    double pos[3], v[3], w[3], R[3][3];
    double root_q[4];
    double angle[NJ], rate[NJ], torque[NJ];
    double contact[NCH];
    int flags;
};

__device__ void task_view(const DevWorlds &d, const Model &M, size_t i, TaskView &t)
{
    const size_t S = d.stride;
    Q4 q = q4(d.q[0 * S + i], d.q[1 * S + i], d.q[2 * S + i], d.q[3 * S + i]);
    M3 R;
    m_from_q(R, q);
    for (int a = 0; a < 3; ++a) {
        t.pos[a] = d.x[a * S + i];
        t.v[a] = d.v[a * S + i];
        t.w[a] = d.w[a * S + i];
        for (int c = 0; c < 3; ++c)
            t.R[a][c] = R.e[a][c];
    }
    t.root_q[0] = q.w; t.root_q[1] = q.x; t.root_q[2] = q.y; t.root_q[3] = q.z;
    for (int j = 0; j < NJ; ++j) {
        t.angle[j] = d.jout[(0 * NJ + j) * S + i];
        t.rate[j] = d.jout[(1 * NJ + j) * S + i];
        t.torque[j] = d.jout[(2 * NJ + j) * S + i];
    }
    float o[NCH];
    for (int c = 0; c < NCH; ++c)
        o[c] = 0.0f;
    static_for<0, NS>([&](auto s) {
        constexpr int ch = sphere_channel(s);
        o[ch] = maxf_(o[ch], (float)d.touch[s * S + i] * M.inv_sub);
    });
    for (int c = 0; c < NCH; ++c)
        t.contact[c] = o[c];
    t.flags = d.unstable[i];
}

/* unsigned angle between the gaze (body +x on the ground) and the target,
 * arc/task.py:80-87 */
__device__ double gaze_error(const double R[3][3], const double *pos, const double *target)
{
    const double f0 = R[0][0], f1 = R[1][0];
    const double d0 = target[0] - pos[0], d1 = target[1] - pos[1];
    const double cross = f0 * d1 - f1 * d0;
    return fabs(atan2(cross, f0 * d0 + f1 * d1));
}

__device__ double task_uniform(TaskDev &td, size_t S, size_t i, double lo, double hi)
{
    Rng r;
    for (int k = 0; k < 4; ++k)
        r.s[k] = td.rng[k * S + i];
    const uint32_t u = r.next();
    for (int k = 0; k < 4; ++k)
        td.rng[k * S + i] = r.s[k];
    return lo + (hi - lo) * ((double)u * 2.3283064365386963e-10);
}

/* world i's position and rotation in binary64, as observe() reports them */
__device__ void task_pose(const DevWorlds &d, size_t i, double pos[3], double R[3][3])
{
    const size_t S = d.stride;
    Q4 q = q4(d.q[0 * S + i], d.q[1 * S + i], d.q[2 * S + i], d.q[3 * S + i]);
    M3 m;
    m_from_q(m, q);
    for (int a = 0; a < 3; ++a) {
        pos[a] = d.x[a * S + i];
        for (int c = 0; c < 3; ++c)
            R[a][c] = m.e[a][c];
    }
}

/* a placed checkpoint and what follows from it, arc/task.py:145-150 */
__device__ void set_target(const DevWorlds &d, const Model &M, TaskDev td, const TaskCfg &c, size_t i,
                           const double pos[3], const double R[3][3], double tx, double ty)
{
    const size_t S = d.stride;
    td.target[0 * S + i] = tx;
    td.target[1 * S + i] = ty;
    TerrainRef tr = terrain_ref(d, i);
    tr.base = d.tbase[i];
    tr.outside = d.outside[i];
    td.target_z[i] = terrain_height(M.terr, tr, (float)tx, (float)ty);
    const double tg[2] = { tx, ty };
    td.prev_dist[i] = hypot(tx - pos[0], ty - pos[1]);
    const double g = gaze_error(R, pos, tg);
    td.prev_gaze[i] = g;
    td.in_view[i] = g < c.fov_rad;
}

/* the next checkpoint, 3-6 m away inside the terrain, arc/task.py:127-150,
 * from world i's own stream */
__device__ void new_target(const DevWorlds &d, const Model &M, TaskDev td, const TaskCfg &c, size_t i,
                           const TaskView &t)
{
    const size_t S = d.stride;
    const double limit = c.half_extent - c.border;
    double tx = 0.0, ty = 0.0;
    bool placed = false;
    for (int k = 0; k < 20 && !placed; ++k) {
        const double ang = task_uniform(td, S, i, -3.141592653589793, 3.141592653589793);
        const double dist = task_uniform(td, S, i, c.checkpoint_min_dist, c.checkpoint_max_dist);
        const double cx = t.pos[0] + dist * cos(ang), cy = t.pos[1] + dist * sin(ang);
        if (fabs(cx) < limit && fabs(cy) < limit) {
            tx = cx;
            ty = cy;
            placed = true;
        }
    }
    if (!placed) { /* towards the centre */
        tx = 0.5 * t.pos[0];
        ty = 0.5 * t.pos[1];
    }
    set_target(d, M, td, c, i, t.pos, t.R, tx, ty);
}

/* drive targets from the actions: default + scale * clip(a, -1, 1) */
__global__ void k_task_actions(DevWorlds d, TaskDev td, TaskCfg c, const float *__restrict__ actions,
                               float *__restrict__ targets)
{
    const size_t S = d.stride;
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)d.n)
        return;
    for (int j = 0; j < NJ; ++j) {
        double a = (double)actions[i * NJ + j];
        a = a < -1.0 ? -1.0 : (a > 1.0 ? 1.0 : a);
        td.action[j * S + i] = a;
        targets[i * NJ + j] = (float)(c.def[j] + c.action_scale * a);
    }
}

/* rewards, terminations, checkpoints and the curriculum after a step,
 * arc/task.py:155-239; ended worlds go to the reset list */
__global__ void k_task_post(DevWorlds d, const Model *__restrict__ m, TaskDev td, TaskCfg c, TaskOut out)
{
    const Model &M = *m;
    const size_t S = d.stride;
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)d.n)
        return;
    TaskView t;
    task_view(d, M, i, t);
    const int steps = td.steps[i] + 1;
    td.steps[i] = steps;
    const double tg[2] = { td.target[0 * S + i], td.target[1 * S + i] };
    const double dx = tg[0] - t.pos[0], dy = tg[1] - t.pos[1];
    const double dist = hypot(dx, dy);
    const double gaze = gaze_error(t.R, t.pos, tg);
    double view = (cos(gaze) - c.cos_fov) / (1.0 - c.cos_fov);
    view = view < 0.0 ? 0.0 : (view > 1.0 ? 1.0 : view);
    const bool in_view = gaze < c.fov_rad;
    double progress = td.prev_dist[i] - dist;
    const double cap = c.max_speed * c.dt;
    progress = progress < cap ? progress : cap;
    if (progress > 0.0)
        progress = progress * view;
    const double turn = td.prev_gaze[i] - gaze;
    const bool reached = dist < c.checkpoint_radius && in_view;

    double br[2];
    for (int a = 0; a < 2; ++a)
        br[a] = t.w[0] * t.R[0][a] + t.w[1] * t.R[1][a] + t.w[2] * t.R[2][a];
    const bool torso = t.contact[LEGS] > 0.0;
    const int tsteps = torso ? td.torso_steps[i] + 1 : 0;
    td.torso_steps[i] = tsteps;
    double power = 0.0, rate = 0.0;
    for (int j = 0; j < NJ; ++j) {
        power += fabs(t.rate[j] * t.torque[j]);
        const double da = td.action[j * S + i] - td.prev_action[j * S + i];
        rate += da * da;
    }
    double reward = c.w_progress * progress + c.w_turn * turn + c.w_checkpoint * (reached ? 1.0 : 0.0) - c.w_time
                    - c.w_energy * power * c.dt - c.w_action_rate * rate
                    - c.w_body_contact * (t.contact[LEGS + 1] > 0.0 ? 1.0 : 0.0)
                    - c.w_torso_contact * (torso ? 1.0 : 0.0) - c.w_body_rate * (br[0] * br[0] + br[1] * br[1])
                    - c.w_bounce * (t.v[2] * t.v[2]);

    bool finite = true;
    for (int a = 0; a < 3; ++a)
        finite = finite && isfinite(t.pos[a]) && isfinite(t.v[a]) && isfinite(t.w[a]);
    for (int a = 0; a < 4; ++a)
        finite = finite && isfinite(t.root_q[a]);
    const bool bad = t.flags != 0 || !finite;
    const bool vd = fabs(t.pos[0]) > c.half_extent || fabs(t.pos[1]) > c.half_extent || t.pos[2] < c.void_z;
    const bool flip = t.R[2][2] < c.flip_up_z;
    const bool beached = tsteps > c.beached_steps;
    const bool fell = flip || beached || vd || bad;
    const int cause = bad ? 1 : vd ? 2 : flip ? 3 : beached ? 4 : 0;
    if (fell)
        reward = reward - c.w_fall;
    if (bad)
        reward = -c.w_fall;

    td.prev_dist[i] = dist;
    td.prev_gaze[i] = gaze;
    td.in_view[i] = in_view;
    for (int j = 0; j < NJ; ++j)
        td.prev_action[j * S + i] = td.action[j * S + i];
    const int cps = td.checkpoints[i] + (reached ? 1 : 0);
    td.checkpoints[i] = cps;
    if (reached && !fell && !c.np_stream)
        new_target(d, M, td, c, i, t);

    const bool truncated = steps >= c.max_steps && !fell;
    out.reward[i] = (float)reward;
    out.terminated[i] = fell ? 1 : 0;
    out.truncated[i] = truncated ? 1 : 0;
    out.cause[i] = cause;
    out.checkpoints[i] = cps;
    out.steps[i] = steps;
    out.difficulty[i] = td.difficulty[i];
    if (fell || truncated) {
        if (c.curriculum) { /* arc/task.py:241-246 */
            const double up = cps >= c.level_up_checkpoints ? 1.0 : 0.0;
            const double down = (fell && cps == 0 && steps < c.level_down_steps) ? 1.0 : 0.0;
            double lvl = td.level[i] + c.level_step * (up - down);
            lvl = rint(lvl * 1e6) / 1e6;
            td.level[i] = lvl < 0.0 ? 0.0 : (lvl > c.max_difficulty ? c.max_difficulty : lvl);
        }
        if (!c.np_stream) {
            const int slot = atomicAdd(td.count, 1);
            td.list[slot] = (int)i;
        }
    }
    if (c.np_stream) /* the ordered lists are built by k_np_hits */
        td.flag[i] = (reached && !fell ? 1 : 0) | (fell || truncated ? 2 : 0);
}

/* seeds and difficulties of the listed worlds' resets; each world draws its
 * seed from its own stream, so the list's order does not matter */
__global__ void k_task_reset_args(DevWorlds d, TaskDev td)
{
    const size_t S = d.stride;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *td.count)
        return;
    const int i = td.list[k];
    Rng r;
    for (int a = 0; a < 4; ++a)
        r.s[a] = td.rng[a * S + i];
    td.seeds[k] = r.next();
    for (int a = 0; a < 4; ++a)
        td.rng[a * S + i] = r.s[a];
    td.difficulty[i] = td.level[i];
    td.diff[k] = (float)td.level[i];
}

/* episode bookkeeping of the listed worlds after their reset, and their first
 * checkpoint, arc/task.py:108-125 */
__global__ void k_task_after_reset(DevWorlds d, const Model *__restrict__ m, TaskDev td, TaskCfg c)
{
    const size_t S = d.stride;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= *td.count)
        return;
    const int i = td.list[k];
    for (int j = 0; j < NJ; ++j)
        td.prev_action[j * S + i] = 0.0;
    td.torso_steps[i] = 0;
    td.steps[i] = 0;
    td.checkpoints[i] = 0;
    if (c.np_stream) /* placed by k_np_place */
        return;
    TaskView t;
    task_view(d, *m, (size_t)i, t);
    new_target(d, *m, td, c, (size_t)i, t);
}

/* the observation vector of world i (arc/task.py:250-280): gravity, body
 * velocities, joint state, previous action, checkpoint in the heading frame,
 * terrain heights around the torso, foot contacts; float32, non-finite -> 0.
 * OBS_GROUP threads per world: the group splits the height samples (each
 * thread computes the torso's yaw frame itself, the same bits in every one),
 * its first thread writes the other 52 values. */
constexpr int OBS_GROUP = 16;
__global__ void k_task_observe(DevWorlds d, const Model *__restrict__ m, TaskDev td, TaskCfg c,
                               const double *__restrict__ offsets, int k_off, float *__restrict__ obs,
                               const int *__restrict__ only, const int *__restrict__ count)
{
    const Model &M = *m;
    const size_t S = d.stride;
    const int tg = blockIdx.x * blockDim.x + threadIdx.x;
    const int t0 = tg / OBS_GROUP, g = tg % OBS_GROUP;
    size_t i;
    if (only) {
        if (t0 >= *count)
            return;
        i = (size_t)only[t0];
    } else {
        if (t0 >= d.n)
            return;
        i = (size_t)t0;
    }
    const int size = 3 + 3 + 3 + 3 * NJ + 3 + k_off + LEGS;
    float *o = obs + i * size;
    auto put = [&](int p, double v) { float f = (float)v; o[p] = isfinite(f) ? f : 0.0f; };
    const int hp = 3 + 3 + 3 + 3 * NJ + 3; /* the first height's position */
    {
        /* heights as arc_sim_heights (csrc/arc_sim.c:185-203): binary32 yaw frame */
        Q4 q = q4(d.q[0 * S + i], d.q[1 * S + i], d.q[2 * S + i], d.q[3 * S + i]);
        V3 f3 = qrot(q, v3(1.0f, 0.0f, 0.0f));
        float yaw = f_atan2(f3.y, f3.x), cy = f_cos(yaw), sy = f_sin(yaw);
        float px = d.x[0 * S + i], py = d.x[1 * S + i];
        const double pz = d.x[2 * S + i];
        TerrainRef tr = terrain_ref(d, i);
        tr.base = d.tbase[i];
        tr.outside = d.outside[i];
        for (int k = g; k < k_off; k += OBS_GROUP) {
            float f = (float)offsets[2 * k], l = (float)offsets[2 * k + 1];
            double h = terrain_height(M.terr, tr, px + cy * f - sy * l, py + sy * f + cy * l);
            double r = pz - c.stand - h;
            put(hp + k, r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r));
        }
    }
    if (g != 0)
        return;
    TaskView t;
    task_view(d, M, i, t);
    int p = 0;
    for (int a = 0; a < 3; ++a)
        put(p++, -t.R[2][a]);
    for (int a = 0; a < 3; ++a)
        put(p++, t.v[0] * t.R[0][a] + t.v[1] * t.R[1][a] + t.v[2] * t.R[2][a]);
    for (int a = 0; a < 3; ++a)
        put(p++, 0.25 * (t.w[0] * t.R[0][a] + t.w[1] * t.R[1][a] + t.w[2] * t.R[2][a]));
    for (int j = 0; j < NJ; ++j)
        put(p++, t.angle[j] - c.def[j]);
    for (int j = 0; j < NJ; ++j)
        put(p++, 0.1 * t.rate[j]);
    for (int j = 0; j < NJ; ++j)
        put(p++, td.prev_action[j * S + i]);
    {
        const double heading = atan2(t.R[1][0], t.R[0][0]);
        const double dx = td.target[0 * S + i] - t.pos[0], dy = td.target[1 * S + i] - t.pos[1];
        const double cs = cos(heading), sn = sin(heading);
        const double fwd = cs * dx + sn * dy, left = -sn * dx + cs * dy;
        double dist = hypot(fwd, left);
        dist = dist > 1e-6 ? dist : 1e-6;
        put(p++, fwd / dist);
        put(p++, left / dist);
        put(p++, (dist < 6.0 ? dist : 6.0) / 6.0);
    }
    p += k_off;
    for (int l = 0; l < LEGS; ++l)
        put(p++, t.contact[l]);
}

__global__ void k_task_init(DevWorlds d, TaskDev td, uint32_t seed, double level)
{
    const size_t S = d.stride;
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)d.n)
        return;
    Rng r;
    r.seed(seed ^ (0x9e3779b9u * (uint32_t)(i + 1)));
    for (int a = 0; a < 4; ++a)
        td.rng[a * S + i] = r.s[a];
    td.level[i] = level;
    td.difficulty[i] = level;
    td.steps[i] = 0;
    td.torso_steps[i] = 0;
    td.checkpoints[i] = 0;
    for (int j = 0; j < NJ; ++j) {
        td.prev_action[j * S + i] = 0.0;
        td.action[j * S + i] = 0.0;
    }
}

__global__ void k_task_list_all(DevWorlds d, TaskDev td)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0)
        *td.count = d.n;
    if (i < d.n)
        td.list[i] = i;
}

/* the rows of the listed worlds, obs -> terminal observations */
__global__ void k_task_copy_listed(TaskDev td, const float *__restrict__ obs, float *__restrict__ dst, int size)
{
    const int k = blockIdx.x;
    if (k >= *td.count)
        return;
    const size_t i = (size_t)td.list[k];
    for (int a = threadIdx.x; a < size; a += blockDim.x)
        dst[i * size + a] = obs[i * size + a];
}

/* ---- numpy-generator mode: one workgroup consumes the stream in order ---- */

enum { NPW = 1024 };

/* rank of this thread among the workgroup's threads with `flag` set, in
 * thread order; the count of all of them in `total` */
__device__ int wg_rank(bool flag, int *lds, int &total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(flag);
    const int in_wave = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0)
        lds[wave] = __popcll(m);
    __syncthreads();
    int before = 0, all = 0;
    for (int w = 0; w < NPW / 64; ++w) {
        const int c = lds[w];
        before += w < wave ? c : 0;
        all += c;
    }
    __syncthreads();
    total = all;
    return before + in_wave;
}

/* checkpoints of the worlds list[0..m), arc/task.py:127-150 with the
 * generator at state s: each round draws an angle for every pending world
 * in list order, then a distance for every pending world; a candidate inside
 * the border is taken, the others stay pending (in order) for the next of at
 * most 20 rounds; the rest go halfway towards the centre.  Called by the
 * whole workgroup; s is the generator state after the draws. */
__device__ void np_place(const DevWorlds &d, const Model &M, TaskDev td, const TaskCfg &c, const int *list, int m,
                         u128 &s, int *lds)
{
    const NpStream &g = *td.np;
    const double limit = c.half_extent - c.border;
    const double pi = 3.141592653589793;
    const double dlo = c.checkpoint_min_dist, dspan = c.checkpoint_max_dist - c.checkpoint_min_dist;
    const int *src = list;
    int *dst = td.wa;
    int K = m;
    for (int round = 0; round < 20 && K > 0; ++round) {
        int next = 0;
        for (int k0 = 0; k0 < K; k0 += NPW) {
            const int k = k0 + (int)threadIdx.x;
            bool pending = false;
            int i = 0;
            if (k < K) {
                i = src[k];
                double pos[3], R[3][3];
                task_pose(d, (size_t)i, pos, R);
                const double ang = -pi + (pi - -pi) * pcg_unit(np_draw(g, s, (uint64_t)k));
                const double dist = dlo + dspan * pcg_unit(np_draw(g, s, (uint64_t)K + (uint64_t)k));
                const double cx = pos[0] + dist * cos(ang), cy = pos[1] + dist * sin(ang);
                if (fabs(cx) < limit && fabs(cy) < limit)
                    set_target(d, M, td, c, (size_t)i, pos, R, cx, cy);
                else
                    pending = true;
            }
            int total;
            const int r = wg_rank(pending, lds, total);
            if (pending)
                dst[next + r] = i;
            next += total;
        }
        s = np_advance(g, s, 2 * (uint64_t)K);
        K = next;
        src = dst;
        dst = dst == td.wa ? td.wb : td.wa;
        __syncthreads(); /* the pending list is complete before the next round reads it */
    }
    for (int k = (int)threadIdx.x; k < K; k += NPW) {
        const int i = src[k];
        double pos[3], R[3][3];
        task_pose(d, (size_t)i, pos, R);
        set_target(d, M, td, c, (size_t)i, pos, R, 0.5 * pos[0], 0.5 * pos[1]);
    }
}

/* after k_task_post: the hit and ended worlds in world order (np.flatnonzero),
 * the hits' new checkpoints (arc/task.py:232-234), the ended worlds as the
 * reset list */
__global__ __launch_bounds__(NPW) void k_np_hits(DevWorlds d, const Model *__restrict__ m, TaskDev td, TaskCfg c)
{
    __shared__ int lds[NPW / 64];
    const int N = d.n;
    int nh = 0, ne = 0;
    for (int i0 = 0; i0 < N; i0 += NPW) {
        const int i = i0 + (int)threadIdx.x;
        const int f = i < N ? td.flag[i] : 0;
        int th, te;
        const int rh = wg_rank((f & 1) != 0, lds, th);
        const int re = wg_rank((f & 2) != 0, lds, te);
        if (f & 1)
            td.hits[nh + rh] = i;
        if (f & 2)
            td.list[ne + re] = i;
        nh += th;
        ne += te;
    }
    if (threadIdx.x == 0)
        *td.count = ne;
    __syncthreads();
    u128 s = td.np->state;
    np_place(d, *m, td, c, td.hits, nh, s, lds);
    __syncthreads();
    if (threadIdx.x == 0)
        td.np->state = s;
}

/* seeds of the listed worlds' resets, rng.integers(0, 2**32, n, uint32)
 * (arc/task.py:117): 32-bit halves of 64-bit outputs, low first, starting with
 * a half the generator kept from an earlier call; and their difficulties */
__global__ __launch_bounds__(NPW) void k_np_seeds(TaskDev td)
{
    NpStream &g = *td.np;
    const int n = *td.count;
    const u128 s = g.state;
    const int has = (int)g.has_uint32;
    const uint32_t kept = g.uinteger;
    for (int k = (int)threadIdx.x; k < n; k += NPW) {
        uint32_t v;
        if (k < has) {
            v = kept;
        } else {
            const int t = k - has;
            const uint64_t x = np_draw(g, s, (uint64_t)(t >> 1));
            v = (t & 1) ? (uint32_t)(x >> 32) : (uint32_t)x;
        }
        td.seeds[k] = v;
        const int i = td.list[k];
        td.difficulty[i] = td.level[i];
        td.diff[k] = (float)td.level[i];
    }
    __syncthreads();
    if (threadIdx.x == 0 && n > 0) {
        const int r = n - has; /* halves taken from new outputs */
        if (r == 0) {
            g.has_uint32 = 0u;
        } else {
            const uint64_t used = (uint64_t)(r + 1) / 2;
            g.uinteger = (uint32_t)(np_draw(g, s, used - 1) >> 32);
            g.state = np_advance(g, s, used);
            g.has_uint32 = (uint32_t)(r & 1);
        }
    }
}

/* first checkpoints of the listed (just reset) worlds, arc/task.py:125 */
__global__ __launch_bounds__(NPW) void k_np_place(DevWorlds d, const Model *__restrict__ m, TaskDev td, TaskCfg c)
{
    __shared__ int lds[NPW / 64];
    u128 s = td.np->state;
    np_place(d, *m, td, c, td.list, *td.count, s, lds);
    __syncthreads();
    if (threadIdx.x == 0)
        td.np->state = s;
}

static unsigned blocks_for(int n) { return (unsigned)((n > 0 ? n : 1) + 255) / 256; }

hipError_t task_launch_init(const DevWorlds &d, const TaskDev &td, uint32_t seed, double level, hipStream_t s)
{
    hipLaunchKernelGGL(k_task_init, dim3(blocks_for(d.n)), dim3(256), 0, s, d, td, seed, level);
    return hipGetLastError();
}
hipError_t task_launch_list_all(const DevWorlds &d, const TaskDev &td, hipStream_t s)
{
    hipLaunchKernelGGL(k_task_list_all, dim3(blocks_for(d.n)), dim3(256), 0, s, d, td);
    return hipGetLastError();
}
hipError_t task_launch_actions(const DevWorlds &d, const TaskDev &td, const TaskCfg &c, const float *actions,
                               float *targets, hipStream_t s)
{
    hipLaunchKernelGGL(k_task_actions, dim3(blocks_for(d.n)), dim3(256), 0, s, d, td, c, actions, targets);
    return hipGetLastError();
}
hipError_t task_launch_post(const DevWorlds &d, const Model *m, const TaskDev &td, const TaskCfg &c, const TaskOut &o,
                            hipStream_t s)
{
    hipLaunchKernelGGL(k_task_post, dim3(blocks_for(d.n)), dim3(256), 0, s, d, m, td, c, o);
    return hipGetLastError();
}
hipError_t task_launch_copy_listed(const DevWorlds &d, const TaskDev &td, const float *obs, float *dst, int size,
                                   hipStream_t s)
{
    hipLaunchKernelGGL(k_task_copy_listed, dim3((unsigned)(d.n > 0 ? d.n : 1)), dim3(64), 0, s, td, obs, dst, size);
    return hipGetLastError();
}
hipError_t task_launch_reset_args(const DevWorlds &d, const TaskDev &td, hipStream_t s)
{
    hipLaunchKernelGGL(k_task_reset_args, dim3(blocks_for(d.n)), dim3(256), 0, s, d, td);
    return hipGetLastError();
}
hipError_t task_launch_after_reset(const DevWorlds &d, const Model *m, const TaskDev &td, const TaskCfg &c,
                                   hipStream_t s)
{
    hipLaunchKernelGGL(k_task_after_reset, dim3(blocks_for(d.n)), dim3(256), 0, s, d, m, td, c);
    return hipGetLastError();
}
hipError_t task_launch_np_hits(const DevWorlds &d, const Model *m, const TaskDev &td, const TaskCfg &c,
                               hipStream_t s)
{
    hipLaunchKernelGGL(k_np_hits, dim3(1), dim3(NPW), 0, s, d, m, td, c);
    return hipGetLastError();
}
hipError_t task_launch_np_seeds(const TaskDev &td, hipStream_t s)
{
    hipLaunchKernelGGL(k_np_seeds, dim3(1), dim3(NPW), 0, s, td);
    return hipGetLastError();
}
hipError_t task_launch_np_place(const DevWorlds &d, const Model *m, const TaskDev &td, const TaskCfg &c,
                                hipStream_t s)
{
    hipLaunchKernelGGL(k_np_place, dim3(1), dim3(NPW), 0, s, d, m, td, c);
    return hipGetLastError();
}
hipError_t task_launch_observe(const DevWorlds &d, const Model *m, const TaskDev &td, const TaskCfg &c,
                               const double *offsets, int k_off, float *obs, bool listed, hipStream_t s)
{
    hipLaunchKernelGGL(k_task_observe, dim3(blocks_for(d.n * OBS_GROUP)), dim3(256), 0, s, d, m, td, c, offsets, k_off, obs,
                       listed ? td.list : nullptr, listed ? td.count : nullptr);
    return hipGetLastError();
}

} /* namespace xpbd */
