/*
 * wave_tables.cpp -- host construction of WaveK (see wave_tables.h).
 */
#include "wave_tables.h"
#include <string.h>
#include <algorithm>
#include "model.h"

namespace xpbd {

bool build_wave_tables(const HostWorld &w, float cell, WaveK &T)
{
    memset(&T, 0, sizeof(T));
    if (w.nb > WAVE_MAX_BODIES || w.nj > WAVE_MAX_JOINTS || w.ns > WAVE_MAX_SPHERES)
        return false;
    T.nb = w.nb;
    T.nj = w.nj;
    T.ns = w.ns;
    /* level of a joint: one more than the deepest earlier joint it shares a body with */
    int level[WAVE_MAX_JOINTS];
    T.nlev = 0;
    for (int j = 0; j < w.nj; ++j) {
        T.jp[j] = w.j[j].parent;
        T.jc[j] = w.j[j].child;
        level[j] = 0;
        for (int k = 0; k < j; ++k) {
            bool share = w.j[k].child == w.j[j].child || w.j[k].child == w.j[j].parent
                         || (w.j[k].parent >= 0 && (w.j[k].parent == w.j[j].child || w.j[k].parent == w.j[j].parent));
            if (share)
                level[j] = std::max(level[j], level[k] + 1);
        }
        T.nlev = std::max(T.nlev, level[j] + 1);
    }
    int pos = 0;
    for (int L = 0; L < T.nlev; ++L) {
        T.lev_first[L] = pos;
        for (int j = 0; j < w.nj; ++j)
            if (level[j] == L)
                T.lev_joint[pos++] = j;
    }
    T.lev_first[T.nlev] = pos;
    /* contact groups: bodies with spheres, in order of first sphere */
    int slot = 0, worst = 0, sp = 0;
    T.ncb = 0;
    for (int s = 0; s < w.ns; ++s) {
        int B = w.s[s].body;
        bool seen = false;
        for (int g = 0; g < T.ncb; ++g)
            seen |= T.cb_body[g] == B;
        if (seen)
            continue;
        int g = T.ncb++;
        T.cb_body[g] = B;
        T.cb_first[g] = sp;
        int cap = 0;
        for (int k = s; k < w.ns; ++k)
            if (w.s[k].body == B) {
                T.cb_sphere[sp++] = k;
                T.s_group[k] = g;
                cap += window_cells(w.s[k].rad, cell);
            }
        worst += cap;
        cap = std::min(cap, MAXC);
        T.cb_slot[g] = slot;
        T.cb_cap[g] = cap;
        slot += cap;
    }
    T.cb_first[T.ncb] = sp;
    /* where every window's worst case fits the cap, each group gets its own
     * slot range (at most MAXC in all); otherwise the records share one list
     * of MAXC, in the world's sphere order */
    T.serial_collide = worst > MAXC;
    static_assert(WAVE_MAX_SLOTS >= MAXC, "the serial list holds the reference's cap");
    return T.serial_collide || slot <= WAVE_MAX_SLOTS;
}

} /* namespace xpbd */
