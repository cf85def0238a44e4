/*
 * pcg64.h -- numpy's default generator, np.random.default_rng(seed), on the
 * host and the device, for the task's checkpoint placement and reset seeds
 * (arc/task.py:93,117,136-137 of the reference draw from one such generator).
 *
 *  - seeding: numpy's SeedSequence (pool of four 32-bit words, the hash-mix
 *    of M. O'Neill's seed_seq_fe) expanded to four 64-bit words, which set
 *    PCG64's state and increment by pcg_setseq_128_srandom_r;
 *  - PCG64: the 128-bit LCG s <- s * 0x2360ED051FC65DA44385DF649FCCF645 + inc
 *    with the XSL-RR output rotr64(hi ^ lo, hi >> 58);
 *  - Generator.integers(0, 2**32, dtype=uint32): 32-bit draws, the low half of
 *    a 64-bit output first, the high half kept in the generator for the next
 *    32-bit draw (has_uint32 / uinteger, untouched by 64-bit draws);
 *  - Generator.uniform(lo, hi): lo + (hi - lo) * ((x >> 11) * 2^-53);
 *  - jump-ahead: k steps of the LCG are the affine map s -> A_k s + C_k mod
 *    2^128; the maps of 2^b steps are tabulated once per increment (F. Brown,
 *    "Random number generation with arbitrary strides", 1994), so the draw at
 *    any offset costs at most one multiply-add per set bit of the offset.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xpbd {

typedef unsigned __int128 u128;

#define PD __host__ __device__ __forceinline__

PD u128 pcg_mult() { return ((u128)0x2360ED051FC65DA4ull << 64) | (u128)0x4385DF649FCCF645ull; }
PD uint64_t pcg_output(u128 s)
{
    const uint64_t hi = (uint64_t)(s >> 64), x = hi ^ (uint64_t)s;
    const unsigned r = (unsigned)(hi >> 58);
    return (x >> r) | (x << ((64u - r) & 63u));
}
PD double pcg_unit(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

/* s -> a s + c */
struct PcgAffine {
    u128 a, c;
};
/* `first`, then `then` */
PD PcgAffine pcg_compose(const PcgAffine &first, const PcgAffine &then)
{
    PcgAffine r;
    r.a = then.a * first.a;
    r.c = then.a * first.c + then.c;
    return r;
}

enum { NP_JUMP_BITS = 40 };

/* the generator's state, numpy's 32-bit buffer, and the maps of 2^b steps */
struct NpStream {
    u128 state, inc;
    uint32_t has_uint32, uinteger;
    uint32_t pad[2];
    PcgAffine pow2[NP_JUMP_BITS];
};

/* the state k steps after s (k < 2^40) */
PD u128 np_advance(const NpStream &g, u128 s, uint64_t k)
{
    for (int b = 0; b < NP_JUMP_BITS && k; ++b, k >>= 1)
        if (k & 1)
            s = g.pow2[b].a * s + g.pow2[b].c;
    return s;
}
/* the k-th 64-bit output (0-based) of the generator at state s */
PD uint64_t np_draw(const NpStream &g, u128 s, uint64_t k) { return pcg_output(np_advance(g, s, k + 1)); }

/* numpy's SeedSequence(entropy).generate_state(4, uint64) for an entropy
 * given as its little-endian 32-bit words, then PCG64's seeding */
inline void np_seed(NpStream &g, const uint32_t *words, int n_words)
{
    const uint32_t zero = 0u;
    if (n_words <= 0) {
        words = &zero;
        n_words = 1;
    }
    uint32_t hc = 0x43b0d7e5u;
    auto hashmix = [&hc](uint32_t v) {
        v ^= hc;
        hc *= 0x931e8875u;
        v *= hc;
        return v ^ (v >> 16);
    };
    auto mix = [](uint32_t x, uint32_t y) {
        uint32_t r = 0xca01f9ddu * x - 0x4973f715u * y;
        return r ^ (r >> 16);
    };
    uint32_t pool[4];
    for (int i = 0; i < 4; ++i)
        pool[i] = hashmix(i < n_words ? words[i] : 0u);
    for (int src = 0; src < 4; ++src)
        for (int dst = 0; dst < 4; ++dst)
            if (src != dst)
                pool[dst] = mix(pool[dst], hashmix(pool[src]));
    for (int src = 4; src < n_words; ++src)
        for (int dst = 0; dst < 4; ++dst)
            pool[dst] = mix(pool[dst], hashmix(words[src]));
    uint32_t out[8];
    uint32_t hb = 0x8b51f9ddu;
    for (int i = 0; i < 8; ++i) {
        uint32_t v = pool[i & 3] ^ hb;
        hb *= 0x58f38dedu;
        v *= hb;
        out[i] = v ^ (v >> 16);
    }
    uint64_t w[4];
    for (int k = 0; k < 4; ++k)
        w[k] = (uint64_t)out[2 * k] | ((uint64_t)out[2 * k + 1] << 32);
    const u128 init_state = ((u128)w[0] << 64) | w[1];
    const u128 init_seq = ((u128)w[2] << 64) | w[3];
    g.inc = (init_seq << 1) | 1u;
    g.state = 0;
    g.state = g.state * pcg_mult() + g.inc;
    g.state += init_state;
    g.state = g.state * pcg_mult() + g.inc;
    g.has_uint32 = 0u;
    g.uinteger = 0u;
    g.pad[0] = g.pad[1] = 0u;
    PcgAffine p;
    p.a = pcg_mult();
    p.c = g.inc;
    for (int b = 0; b < NP_JUMP_BITS; ++b) {
        g.pow2[b] = p;
        p = pcg_compose(p, p);
    }
}

#undef PD

} /* namespace xpbd */
