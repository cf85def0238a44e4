/*
 * hostcheck.cpp -- test harness: the device engine's own code (lane_engine.h,
 * reset_engine.h, setup.cpp, model.cpp) compiled for the host CPU and driven
 * like the arc_sim API, so the parity tests can compare it with the oracle
 * without a GPU.  The only difference from the device build is the libm used
 * for the binary64 atan2 / sin / cos (glibc here, the device library there).
 * Never part of the product.
 */
#include <stdlib.h>
#include <string.h>
#include <vector>
#include "pcg64.h"
#include "reset_engine.h"
#include "setup.h"

using namespace xpbd;

struct HC {
    int n;
    SimSetup su;
    std::vector<LaneWorld> w;
    std::vector<ArrayAnchors> anc;
    std::vector<int8_t> tq;
    std::vector<float> tbase, outside;
    int pitch, tcells;
    std::vector<float> cbuf;
    std::vector<int> err;
};

extern "C" {

HC *hc_create(int n, double dt, int substeps, int grid, double cell)
{
    HC *h = new HC();
    h->n = n;
    if (!sim_setup(h->su, dt, substeps, grid, cell)) {
        delete h;
        return nullptr;
    }
    const int g = h->su.grid;
    h->pitch = g + 2 * TPAD;
    h->tcells = h->pitch * h->pitch;
    h->w.resize(n);
    h->anc.assign(n, ArrayAnchors());
    memset(h->anc.data(), 0, sizeof(ArrayAnchors) * (size_t)n);
    h->tq.assign((size_t)n * h->tcells, (int8_t)Q_OUT);
    for (int i = 0; i < n; ++i)
        for (int iy = 0; iy < g; ++iy)
            for (int ix = 0; ix < g; ++ix)
                h->tq[(size_t)i * h->tcells + (ix + TPAD) + (iy + TPAD) * h->pitch] = 0;
    h->tbase.assign(n, 0.0f);
    h->outside.assign(n, h->su.tmpl.outside);
    h->cbuf.assign((size_t)h->su.model.cslots * 6, 0.0f);
    h->err.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        LaneWorld &W = h->w[i];
        memset(&W, 0, sizeof(W));
        for (int b = 0; b < NB; ++b) {
            W.b[b].x = h->su.tk.x[b];
            W.b[b].q = h->su.tk.q[b];
        }
        for (int j = 0; j < NJ; ++j) {
            W.target[j] = h->su.tk.target[j];
            W.angle[j] = h->su.tk.angle[j];
        }
    }
    return h;
}

void hc_destroy(HC *h) { delete h; }

void hc_reset(HC *h, const int *ids, int n_ids, const unsigned *seeds, const double *diff)
{
    const Model &M = h->su.model;
    const ResetK &R = h->su.rk;
    const int n = R.grid, cells = n * n;
    std::vector<float> lat(LAT_LIMIT), hb(cells);
    float draws[NJ + 1];
    for (int k = 0; k < n_ids; ++k) {
        int i = ids[k];
        float d = (float)diff[k];
        bool gen = d > 0.0f;
        reset_draws(seeds[k], gen, n, lat.data(), hb.data(), draws);
        for (int c = 0; c < cells; ++c)
            hb[c] = gen ? column_step(lat.data(), n, d, hb[c], c) : 0.0f;
        float base = 0.0f;
        if (gen) {
            int ix = (int)floorf((0.0f - M.terr.ox) / M.terr.cell), iy = (int)floorf((0.0f - M.terr.oy) / M.terr.cell);
            base = 0.03f * hb[ix + iy * n];
        }
        int8_t *q = &h->tq[(size_t)i * h->tcells];
        for (int c = 0; c < h->tcells; ++c) {
            int ix = c % h->pitch - TPAD, iy = c / h->pitch - TPAD;
            q[c] = (ix >= 0 && iy >= 0 && ix < n && iy < n) ? (int8_t)(int)hb[ix + iy * n] : (int8_t)Q_OUT;
        }
        h->tbase[i] = base;
        h->outside[i] = -100.0f;
        TerrainRef tl;
        tl.q = q + TPAD + TPAD * h->pitch;
        tl.pitch = h->pitch;
        tl.base = base;
        tl.outside = -100.0f;
        float ang[NJ];
        for (int j = 0; j < NJ; ++j)
            ang[j] = clampf_(R.def[j] + draws[j], R.lo[j], R.hi[j]);
        float ground = ground_under(M.terr, tl, R.footprint, R.cell);
        V3 bx[NB];
        Q4 bq[NB];
        pose_bodies(M, v3(0.0f, 0.0f, ground + R.stand + 0.03f), qaxis(v3(0.0f, 0.0f, 1.0f), draws[NJ]), ang, bx, bq);
        LaneWorld &W = h->w[i];
        memset(&W, 0, sizeof(W));
        memset(&h->anc[i], 0, sizeof(ArrayAnchors));
        for (int b = 0; b < NB; ++b) {
            W.b[b].x = bx[b];
            W.b[b].q = bq[b];
        }
        for (int j = 0; j < NJ; ++j) {
            W.target[j] = ang[j];
            W.angle[j] = ang[j];
        }
        h->err[i] = 0;
    }
}

void hc_step(HC *h, const double *targets)
{
    const Model &M = h->su.model;
    const int cells = h->su.grid * h->su.grid;
    for (int i = 0; i < h->n; ++i) {
        LaneWorld &W = h->w[i];
        for (int j = 0; j < NJ; ++j)
            W.target[j] = clampf_((float)targets[i * NJ + j], M.joint[j].lo, M.joint[j].hi);
        if (W.unstable)
            continue;
        TerrainRef tr;
        tr.q = &h->tq[(size_t)i * h->tcells] + TPAD + TPAD * h->pitch;
        tr.pitch = h->pitch;
        tr.base = h->tbase[i];
        tr.outside = h->outside[i];
        ContactStore cs;
        cs.base = h->cbuf.data();
        cs.stride = 1;
        cs.world = 0;
        ArrayPoses poses;
        h->err[i] |= lane_step(W, h->anc[i], poses, M, tr, cs);
    }
}

/* oracle dump layout */
void hc_state(const HC *h, int i, double *out)
{
    const LaneWorld &W = h->w[i];
    int o = 0;
    for (int b = 0; b < NB; ++b) {
        const Body &B = W.b[b];
        double v[13] = { B.x.x, B.x.y, B.x.z, B.q.w, B.q.x, B.q.y, B.q.z, B.v.x, B.v.y, B.v.z, B.w.x, B.w.y, B.w.z };
        for (int c = 0; c < 13; ++c) out[o++] = v[c];
    }
    for (int j = 0; j < NJ; ++j) {
        out[o++] = W.target[j]; out[o++] = W.angle[j]; out[o++] = W.rate[j]; out[o++] = W.torque[j];
    }
    const ArrayAnchors &A = h->anc[i];
    for (int s = 0; s < NS; ++s) {
        out[o++] = touch_of(W, s);
        out[o++] = (W.anchored >> s) & 1u;
        out[o++] = A.a[s].x; out[o++] = A.a[s].y; out[o++] = A.a[s].z;
        out[o++] = A.n[s].x; out[o++] = A.n[s].y; out[o++] = A.n[s].z;
    }
    out[o++] = W.unstable;
    out[o++] = W.time;
}

void hc_terrain(const HC *h, int i, double *out)
{
    const int n = h->su.grid;
    TerrainRef r;
    r.q = &h->tq[(size_t)i * h->tcells] + TPAD + TPAD * h->pitch;
    r.pitch = h->pitch;
    r.base = h->tbase[i];
    r.outside = h->outside[i];
    for (int iy = 0; iy < n; ++iy)
        for (int ix = 0; ix < n; ++ix)
            out[ix + iy * n] = terrain_cell(h->su.model.terr, r, ix, iy);
    out[n * n] = h->outside[i];
}

int hc_error(const HC *h, int i) { return h->err[i]; }

}

extern "C" void hc_joint_consts(const HC *h, int j, double *o)
{
    const JointK &k = h->su.model.joint[j];
    const V3 v[6] = { k.r1, k.r2, k.a1, k.a2, k.b1, k.b2 };
    for (int i = 0; i < 6; ++i) {
        o[3 * i] = v[i].x; o[3 * i + 1] = v[i].y; o[3 * i + 2] = v[i].z;
    }
    o[18] = k.qrel0.w; o[19] = k.qrel0.x; o[20] = k.qrel0.y; o[21] = k.qrel0.z;
}

/* The reset's draws made as RESET_LANES jump-ahead chunks (the device's way) against
 * the one sequential stream, and the split footprint scan against the plain
 * one.  Returns the number of differing values (0 expected). */
/* The column quotient's sequence (xpbd_core.h qdiv_fma, the device's fast
 * path) on the host, whose fmaf is the same correctly rounded fused
 * multiply-add, against the IEEE quotient a / d over the binary32 patterns
 * lo, lo + step, ... below hi (NaN and infinities skipped; with in_domain, also
 * the numerators outside qdiv's domain).  Returns the mismatches. */
extern "C" unsigned long long hc_quotient_mismatches(float d, unsigned long long lo, unsigned long long hi,
                                                     unsigned long long step, int in_domain)
{
    const float y = 1.0f / d;
    unsigned long long bad = 0;
    for (unsigned long long u = lo; u < hi; u += step) {
        const uint32_t bits = (uint32_t)u;
        float a;
        memcpy(&a, &bits, 4);
        if (!(fabsf(a) <= 3.4028235e38f) || (in_domain && !xpbd::qdiv_domain(a)))
            continue;
        const float f = xpbd::qdiv_fma(a, d, y), q = a / d;
        bad += memcmp(&f, &q, 4) != 0 ? 1u : 0u;
    }
    return bad;
}

extern "C" int hc_reset_split_check(const HC *h, unsigned seed, int gen)
{
    const ResetK &R = h->su.rk;
    const Model &M = h->su.model;
    const int n = R.grid, cells = n * n;
    std::vector<float> lat(LAT_LIMIT, 0.0f), jit(cells, 0.0f), lat2(LAT_LIMIT, 0.0f), jit2(cells, 0.0f);
    float d1[NJ + 1], d2[NJ + 1];
    reset_draws(seed, gen != 0, n, lat.data(), jit.data(), d1);
    if (gen)
        for (int lane = 0; lane < RESET_LANES; ++lane)
            reset_draws_lane(seed, true, n, h->su.jtab.data(), R.chunk, lane, lat2.data(), jit2.data(), d2);
    else
        reset_draws(seed, false, n, lat2.data(), jit2.data(), d2);
    int bad = 0;
    for (int c = 0; c < LAT_LIMIT; ++c)
        bad += memcmp(&lat[c], &lat2[c], 4) != 0;
    for (int c = 0; c < cells; ++c)
        bad += memcmp(&jit[c], &jit2[c], 4) != 0;
    for (int c = 0; c <= NJ; ++c)
        bad += memcmp(&d1[c], &d2[c], 4) != 0;
    /* footprint scan over a world's terrain built from these draws */
    std::vector<int8_t> q((size_t)(n + 2 * TPAD) * (n + 2 * TPAD), (int8_t)Q_OUT);
    const int pitch = n + 2 * TPAD;
    for (int c = 0; c < cells; ++c) {
        float F = gen ? column_step(lat.data(), n, 0.7f, jit[c], c) : 0.0f;
        q[(c % n + TPAD) + (c / n + TPAD) * pitch] = (int8_t)(int)F;
    }
    TerrainRef tl;
    tl.q = q.data() + TPAD + TPAD * pitch;
    tl.pitch = pitch;
    tl.base = 0.09f;
    tl.outside = -100.0f;
    float g1 = ground_under(M.terr, tl, R.footprint, R.cell), g2 = -1e30f;
    for (int lane = 0; lane < RESET_LANES; ++lane)
        g2 = maxf_(g2, ground_part(M.terr, tl, R.footprint, R.cell, lane, RESET_LANES));
    bad += memcmp(&g1, &g2, 4) != 0;
    return bad;
}

/* numpy's generator as the device task draws from it (pcg64.h): the seeded
 * state and increment as hi/lo words (out4), and `count` 64-bit outputs at the
 * offsets `ks` from the seeded state, each taken by jump-ahead. */
extern "C" void hc_np_seed(const unsigned *words, int n_words, unsigned long long *out4)
{
    NpStream g;
    np_seed(g, words, n_words);
    out4[0] = (unsigned long long)(g.state >> 64);
    out4[1] = (unsigned long long)g.state;
    out4[2] = (unsigned long long)(g.inc >> 64);
    out4[3] = (unsigned long long)g.inc;
}

extern "C" void hc_np_draws(const unsigned *words, int n_words, const unsigned long long *ks, int count,
                            unsigned long long *out)
{
    NpStream g;
    np_seed(g, words, n_words);
    for (int i = 0; i < count; ++i)
        out[i] = np_draw(g, g.state, ks[i]);
}
