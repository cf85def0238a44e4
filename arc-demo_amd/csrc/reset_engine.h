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

/* xorshift128 is linear over GF(2)^128: the state after k draws is a fixed
 * 128 x 128 bit matrix times the state.  A matrix is kept as the images of
 * the 128 unit states (bit j = word j / 32, bit j % 32), so applying it is an
 * XOR of the columns selected by the state's set bits. */
struct RngMatrix {
    uint32_t col[128][4];
};
XD void rng_apply(const RngMatrix &m, uint32_t s[4])
{
    uint32_t r0 = 0u, r1 = 0u, r2 = 0u, r3 = 0u;
    for (int j = 0; j < 128; ++j) {
        const uint32_t mask = 0u - ((s[j >> 5] >> (j & 31)) & 1u);
        r0 ^= m.col[j][0] & mask;
        r1 ^= m.col[j][1] & mask;
        r2 ^= m.col[j][2] & mask;
        r3 ^= m.col[j][3] & mask;
    }
    s[0] = r0;
    s[1] = r1;
    s[2] = r2;
    s[3] = r3;
}
/* the matrix of one draw, and products of matrices (a after b) */
inline void rng_step_matrix(RngMatrix &m)
{
    for (int j = 0; j < 128; ++j) {
        Rng r;
        r.s[0] = r.s[1] = r.s[2] = r.s[3] = 0u;
        r.s[j >> 5] = 1u << (j & 31);
        r.next();
        for (int w = 0; w < 4; ++w)
            m.col[j][w] = r.s[w];
    }
}
inline void rng_compose(const RngMatrix &a, const RngMatrix &b, RngMatrix &out)
{
    RngMatrix t;
    for (int j = 0; j < 128; ++j) {
        uint32_t v[4] = { b.col[j][0], b.col[j][1], b.col[j][2], b.col[j][3] };
        rng_apply(a, v);
        for (int w = 0; w < 4; ++w)
            t.col[j][w] = v[w];
    }
    out = t;
}
/* the matrix of k draws */
inline void rng_power(uint64_t k, RngMatrix &out)
{
    RngMatrix p, r;
    rng_step_matrix(p);
    bool any = false;
    while (k) {
        if (k & 1u) {
            if (any)
                rng_compose(p, r, r);
            else
                r = p;
            any = true;
        }
        k >>= 1;
        if (k)
            rng_compose(p, p, p);
    }
    if (!any) { /* identity */
        for (int j = 0; j < 128; ++j)
            for (int w = 0; w < 4; ++w)
                r.col[j][w] = (w == (j >> 5)) ? 1u << (j & 31) : 0u;
    }
    out = r;
}

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

/* Draws of a reset spread over the RESET_LANES lanes of its workgroup: lane
 * L makes the draws [L * chunk, (L + 1) * chunk) of the stream, starting from
 * the seeded state moved L * chunk draws ahead.  The generator is linear over
 * GF(2)^128, so that move is one fixed 128 x 128 bit matrix per lane, applied
 * from a table of 4-bit windows: entry [w][v] is the XOR of the matrix columns
 * 4w..4w+3 selected by the bits of v, so a product is 32 lookups.  Writes
 * each draw where reset_draws puts it. */
constexpr int RESET_LANES = 256;
constexpr int JTAB_WORDS = 32 * 16 * 4; /* one lane's table: 32 windows x 16 entries x 4 words */

XD int reset_draw_count(bool gen, int n)
{
    const int m1 = lattice1(n), m2 = lattice2(n);
    return (gen ? m1 * m1 + m2 * m2 + n * n : 0) + NJ + 1;
}
XD int reset_chunk(int n) { return (reset_draw_count(true, n) + RESET_LANES - 1) / RESET_LANES; }

XD void rng_apply_tab(const uint32_t *tab, uint32_t s[4])
{
    uint32_t r0 = 0u, r1 = 0u, r2 = 0u, r3 = 0u;
#pragma unroll
    for (int w = 0; w < 32; ++w) {
        const uint32_t v = (s[w >> 3] >> (4 * (w & 7))) & 15u;
        const uint32_t *e = tab + (w * 16 + (int)v) * 4;
        r0 ^= e[0];
        r1 ^= e[1];
        r2 ^= e[2];
        r3 ^= e[3];
    }
    s[0] = r0;
    s[1] = r1;
    s[2] = r2;
    s[3] = r3;
}

/* the window tables of the matrices of L * chunk draws, L = 0..RESET_LANES-1
 * (RESET_LANES * JTAB_WORDS words) */
inline void rng_lane_tables(int chunk, uint32_t *out)
{
    RngMatrix step, m;
    rng_power((uint64_t)chunk, step);
    rng_power(0, m); /* identity */
    for (int L = 0; L < RESET_LANES; ++L) {
        uint32_t *t = out + (size_t)L * JTAB_WORDS;
        for (int w = 0; w < 32; ++w)
            for (int v = 0; v < 16; ++v) {
                uint32_t e[4] = { 0u, 0u, 0u, 0u };
                for (int b = 0; b < 4; ++b)
                    if ((v >> b) & 1)
                        for (int k = 0; k < 4; ++k)
                            e[k] ^= m.col[4 * w + b][k];
                for (int k = 0; k < 4; ++k)
                    t[(w * 16 + v) * 4 + k] = e[k];
            }
        rng_compose(step, m, m);
    }
}

XD void reset_draws_lane(uint32_t seed, bool gen, int n, const uint32_t *jtab, int chunk, int lane, float *lat,
                         float *jit, float *draws)
{
    const int m1 = lattice1(n), m2 = lattice2(n);
    const int nl = gen ? m1 * m1 + m2 * m2 : 0, nj = gen ? n * n : 0;
    const int total = nl + nj + NJ + 1;
    const int a = lane * chunk;
    if (a >= total)
        return;
    Rng r;
    r.seed(seed);
    if (lane > 0)
        rng_apply_tab(jtab + (size_t)lane * JTAB_WORDS, r.s);
    const int e = a + chunk < total ? a + chunk : total;
    for (int q = a; q < e; ++q) {
        if (q < nl) {
            lat[q] = r.range(-1.0f, 1.0f);
        } else if (q < nl + nj) {
            jit[q - nl] = r.range(-1.0f, 1.0f);
        } else {
            const int k = q - nl - nj;
            draws[k] = k < NJ ? r.range(-0.1f, 0.1f) : r.range(-3.14159265358979323846f, 3.14159265358979323846f);
        }
    }
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

/* the samples of ground_under along one axis: v_0 = -reach, v_k+1 = v_k +
 * half a cell while v <= reach (the reference's float loop) */
XD int ground_samples(float reach, float cell)
{
    int k = 0;
    for (float v = -reach; v <= reach; v += 0.5f * cell)
        ++k;
    return k;
}
XD float ground_sample(float reach, float cell, int k)
{
    float v = -reach;
    for (int i = 0; i < k; ++i)
        v += 0.5f * cell;
    return v;
}
/* the maximum over samples [p0, p0 + step, ...) of the K x K grid (x outer) */
XD float ground_part(const TerrainK &t, const TerrainRef &tr, float footprint, float cell, int p0, int step)
{
    const float reach = footprint + 0.1f;
    const int K = ground_samples(reach, cell);
    float ground = -1e30f;
    for (int p = p0; p < K * K; p += step)
        ground = maxf_(ground, terrain_height(t, tr, ground_sample(reach, cell, p / K), ground_sample(reach, cell, p % K)));
    return ground;
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
