/*
 * reset_engine.h -- the pieces of a world reset (csrc/arc_sim.c:88-110):
 * xorshift128 with hash seeding (csrc/arc_rng.h:14-53), block-terrain
 * generation (csrc/arc_terrain.c:32-71), the footprint ground scan and the
 * forward kinematics of the pose (csrc/arc_phys.c:749-782), in binary32 in
 * the reference's order.  Host and device compile the same code.
 */
#pragma once
#include "lane_engine.h"

namespace xpbd {

struct Rng {
    uint32_t s[4];
    XD void seed(uint32_t v)
    {
        for (int i = 0; i < 4; ++i) {
            uint32_t x = v + 0x9e3779b9u * (uint32_t)(i + 1);
            x = (x ^ (x >> 16)) * 0x7feb352du;
            x = (x ^ (x >> 15)) * 0x846ca68bu;
            x = x ^ (x >> 16);
            v = x;
            s[i] = x ? x : 1u;
        }
    }
    XD uint32_t next()
    {
        uint32_t t = s[3], s0 = s[0];
        s[3] = s[2];
        s[2] = s[1];
        s[1] = s0;
        t ^= t << 11;
        t ^= t >> 8;
        s[0] = t ^ s0 ^ (s0 >> 19);
        return s[0];
    }
    /* lo + (hi - lo) * u, u = float(u32) * 2^-32 (the u32 rounds to binary32 first) */
    XD float range(float lo, float hi) { return lo + (hi - lo) * ((float)next() * 2.3283064365386963e-10f); }
};

/* value-noise lattice sizes for a grid of n columns per side */
XD int lattice1(int n) { return n / 8 + 2; }
XD int lattice2(int n) { return n / 3 + 2; }

XD float smooth3(float t) { return t * t * (3.0f - 2.0f * t); }

/* bilinear value noise with smoothstep weights at column (ix, iy) */
XD float noise_at(const float *lat, int m, int spacing, int ix, int iy)
{
    float fx = (float)ix / (float)spacing, fy = (float)iy / (float)spacing;
    int gx = (int)fx, gy = (int)fy;
    float tx = smooth3(fx - (float)gx), ty = smooth3(fy - (float)gy);
    float a = lat[gx + gy * m], b = lat[gx + 1 + gy * m];
    float c = lat[gx + (gy + 1) * m], e = lat[gx + 1 + (gy + 1) * m];
    return (a + (b - a) * tx) * (1.0f - ty) + (c + (e - c) * tx) * ty;
}

/* Draws of one reset in stream order: lattices, per-column jitter (only for
 * difficulty > 0), then 12 joint-noise draws and the yaw. */
XD void reset_draws(uint32_t seed, bool gen, int n, float *lat, float *jit, float *draws)
{
    Rng r;
    r.seed(seed);
    const int m1 = lattice1(n), m2 = lattice2(n);
    if (gen) {
        for (int c = 0; c < m1 * m1; ++c)
            lat[c] = r.range(-1.0f, 1.0f);
        for (int c = 0; c < m2 * m2; ++c)
            lat[m1 * m1 + c] = r.range(-1.0f, 1.0f);
        for (int c = 0; c < n * n; ++c)
            jit[c] = r.range(-1.0f, 1.0f);
    }
    for (int j = 0; j < NJ; ++j)
        draws[j] = r.range(-0.1f, 0.1f);
    draws[NJ] = r.range(-3.14159265358979323846f, 3.14159265358979323846f);
}

/* quantisation step F of column c before the re-basing: hills + bumps +
 * jitter, h = 0.03f * F with F = floor(h / 0.03 + 0.5) (F is small: |F| <= 21 at
 * difficulty 1) */
XD float column_step(const float *lat, int n, float difficulty, float jitter, int c)
{
    const int m1 = lattice1(n), m2 = lattice2(n);
    int ix = c % n, iy = c / n;
    float h = 0.0f;
    h = h + (0.45f * difficulty) * noise_at(lat, m1, 8, ix, iy);
    h = h + (0.12f * difficulty) * noise_at(lat + m1 * m1, m2, 3, ix, iy);
    h = h + 0.06f * difficulty * jitter;
    return floorf(h / 0.03f + 0.5f);
}

/* highest column under the robot's footprint, csrc/arc_sim.c:104-107 */
XD float ground_under(const TerrainK &t, const TerrainRef &tr, float footprint, float cell)
{
    float reach = footprint + 0.1f, ground = -1e30f;
    for (float x = -reach; x <= reach; x += 0.5f * cell)
        for (float y = -reach; y <= reach; y += 0.5f * cell)
            ground = maxf_(ground, terrain_height(t, tr, x, y));
    return ground;
}

/* forward kinematics of the quadruped from the root pose and joint angles */
XD void pose_bodies(const Model &M, V3 root_x, Q4 root_q, const float *ang, V3 *bx, Q4 *bq)
{
    bx[0] = root_x;
    bq[0] = qunit(root_q);
    for (int j = 0; j < NJ; ++j) {
        const JointK &J = M.joint[j];
        int p = joint_parent(j), c = joint_child(j);
        V3 anchor = add(bx[p], qrot(bq[p], J.r1));
        bq[c] = qunit(qmul(bq[p], qmul(qaxis(J.a1, ang[j]), J.qrel0)));
        bx[c] = sub(anchor, qrot(bq[c], J.r2));
    }
}

} /* namespace xpbd */
