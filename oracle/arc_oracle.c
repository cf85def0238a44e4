/*
 * arc_oracle.c -- CPU restatement of the reference XPBD engine, used ONLY as
 * test infrastructure (the checker for the HIP engine).  Nothing in the
 * product links, loads or calls this file; the tests, the smoke check and the
 * CPU-baseline leg of bench.py do.
 *
 * One source, two builds selected by OREAL:
 *   -DOREAL=double  bit-exact with the reference engine compiled in FP64 with
 *                   -ffp-contract=off (pinned by tests/test_oracle_ref.py and
 *                   the fixtures under tests/golden/);
 *   -DOREAL=float   bit-exact with the reference engine compiled with every
 *                   `double` replaced by `float` and single-precision
 *                   literals (the FP32 build the HIP engine is compared with).
 * Both must be compiled with -ffp-contract=off so that every multiply and add
 * rounds separately, in the order written.
 *
 * Follows, function by function (citations are reference file:line):
 *   csrc/arc_math.h:24-145   vector / quaternion / matrix helpers
 *   csrc/arc_rng.h:14-53     xorshift128 + hash seeding
 *   csrc/arc_terrain.c:17-71 flat / generated block terrain; arc_terrain.h:34-50 lookups
 *   csrc/arc_phys.c          the substep solver (per-function citations below)
 *   csrc/arc_robot.c:72-195  the quadruped (physics part only, no visuals)
 *   csrc/arc_sim.c:26-215    the batched world API
 *   arc/task.py:196-203      fall rules (bad / void / flip / beached)
 *
 * Where the FP32 reference build evaluates a libm result in double before
 * rounding (libm functions keep their double prototypes in that build), the
 * restatement does the same: L1() wraps such calls, and the two mixed
 * expressions of the robot builder are written out in double explicitly.
 *
 * The API at the ctypes boundary is always double; values are rounded to the
 * build's real type on the way in.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <stddef.h>

#ifndef OREAL
#error "compile with -DOREAL=float or -DOREAL=double"
#endif
typedef OREAL real;

#define K(c) ((real)(c))
/* libm in double, result rounded to real (what both reference builds do) */
#define L1(fn, a) ((real)fn((double)(a)))
#define PI_R K(3.14159265358979323846)

enum { NB_MAX = 16, NJ_MAX = 16, NS_MAX = 48, NC_MAX = 128, GRID_MAX = 64 };
enum { LEGS = 4, NJR = 12 };

typedef struct { real x, y, z; } V3;
typedef struct { real w, x, y, z; } Q4;
typedef struct { real e[3][3]; } M33; /* row-major */

typedef struct {
    V3 pos, pos0;          /* centre of mass now / at substep start */
    Q4 rq, rq0;            /* orientation now / at substep start (|q| may drift) */
    V3 lin, ang;           /* world-frame velocities */
    real mass, invm;
    V3 I, invI;            /* principal moments, body frame */
    M33 R;                 /* cached rotation of rq / |rq| */
    M33 Iw;                /* cached R diag(invI) R^T */
} OBody;

typedef struct {
    int parent, child;
    V3 r1, r2, a1, a2, b1, b2;
    Q4 qrel0;
    real lo, hi, kp, kd, tmax, target;
    real angle, rate, torque, tacc;
    real lim[4];
} OJoint;

typedef struct {
    int body;
    V3 at;
    real rad;
    int touch, anchored;
    V3 anc, ancn;
} OSphere;

typedef struct { int body; V3 local, n; } OContact;

typedef struct {
    int nx, ny;
    real cell, ox, oy, outside;
    real h[GRID_MAX * GRID_MAX];
} OTerrain;

typedef struct {
    V3 g;
    real mus, mud;
    int nb, nj, ns, nc;
    int unstable;
    real time;
    OBody b[NB_MAX];
    OJoint j[NJ_MAX];
    OSphere s[NS_MAX];
    OContact c[NC_MAX];
    OTerrain t;
} OWorld;

/* ------------------------------------------------------------------ */
/* small algebra, evaluation order as arc_math.h:24-145                 */

static inline V3 vmk(real x, real y, real z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline V3 vadd(V3 a, V3 b) { return vmk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 vsub(V3 a, V3 b) { return vmk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 vscl(V3 a, real s) { return vmk(a.x * s, a.y * s, a.z * s); }
static inline V3 vneg(V3 a) { return vmk(-a.x, -a.y, -a.z); }
static inline V3 vhad(V3 a, V3 b) { return vmk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline V3 vmad(V3 a, V3 b, real s) { return vmk(a.x + b.x * s, a.y + b.y * s, a.z + b.z * s); }
static inline real vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 vcross(V3 a, V3 b)
{
    return vmk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline real vlen2(V3 a) { return vdot(a, a); }
static inline real vlen(V3 a) { return L1(sqrt, vdot(a, a)); }
static inline V3 vunit(V3 a)
{
    real l = vlen(a);
    return l > K(1e-15) ? vscl(a, K(1.0) / l) : vmk(K(0), K(0), K(0));
}
static inline V3 vperp(V3 n)
{
    V3 t = fabs((double)n.x) < (double)K(0.57) ? vmk(K(1), K(0), K(0)) : vmk(K(0), K(1), K(0));
    return vunit(vcross(n, t));
}

static inline Q4 qmk(real w, real x, real y, real z) { Q4 r; r.w = w; r.x = x; r.y = y; r.z = z; return r; }
static inline Q4 qid(void) { return qmk(K(1), K(0), K(0), K(0)); }
static inline Q4 qconj(Q4 a) { return qmk(a.w, -a.x, -a.y, -a.z); }
static inline Q4 qmul(Q4 a, Q4 b)
{
    return qmk(a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z,
               a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y,
               a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
               a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w);
}
static inline Q4 qunit(Q4 a)
{
    real l = L1(sqrt, a.w * a.w + a.x * a.x + a.y * a.y + a.z * a.z);
    if (l < K(1e-15))
        return qid();
    l = K(1.0) / l;
    return qmk(a.w * l, a.x * l, a.y * l, a.z * l);
}
static inline Q4 qaxis(V3 axis, real angle)
{
    real s = L1(sin, K(0.5) * angle);
    return qmk(L1(cos, K(0.5) * angle), axis.x * s, axis.y * s, axis.z * s);
}
static inline V3 qrot(Q4 q, V3 v)
{
    V3 u = vmk(q.x, q.y, q.z);
    V3 t = vscl(vcross(u, v), K(2.0));
    return vadd(vmad(v, t, q.w), vcross(u, t));
}
static inline V3 qirot(Q4 q, V3 v) { return qrot(qconj(q), v); }

static inline void m_from_q(M33 *r, Q4 q)
{
    real k = K(2.0) / (q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z);
    real x2 = q.x * k, y2 = q.y * k, z2 = q.z * k;
    real xx = q.x * x2, yy = q.y * y2, zz = q.z * z2;
    real xy = q.x * y2, xz = q.x * z2, yz = q.y * z2;
    real wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    r->e[0][0] = K(1.0) - (yy + zz); r->e[0][1] = xy - wz; r->e[0][2] = xz + wy;
    r->e[1][0] = xy + wz; r->e[1][1] = K(1.0) - (xx + zz); r->e[1][2] = yz - wx;
    r->e[2][0] = xz - wy; r->e[2][1] = yz + wx; r->e[2][2] = K(1.0) - (xx + yy);
}
static inline V3 m_mul(const M33 *a, V3 v)
{
    return vmk(a->e[0][0] * v.x + a->e[0][1] * v.y + a->e[0][2] * v.z,
               a->e[1][0] * v.x + a->e[1][1] * v.y + a->e[1][2] * v.z,
               a->e[2][0] * v.x + a->e[2][1] * v.y + a->e[2][2] * v.z);
}
static inline V3 m_tmul(const M33 *a, V3 v)
{
    return vmk(a->e[0][0] * v.x + a->e[1][0] * v.y + a->e[2][0] * v.z,
               a->e[0][1] * v.x + a->e[1][1] * v.y + a->e[2][1] * v.z,
               a->e[0][2] * v.x + a->e[1][2] * v.y + a->e[2][2] * v.z);
}
static inline real rclamp(real v, real lo, real hi) { return v < lo ? lo : (v > hi ? hi : v); }
static inline real rmin(real a, real b) { return a < b ? a : b; }
static inline real rmax(real a, real b) { return a > b ? a : b; }

/* ------------------------------------------------------------------ */
/* RNG, arc_rng.h:14-53 (32-bit words kept in uint32_t)                 */

typedef struct { uint32_t s[4]; } ORng;

static inline uint32_t rng_next(ORng *r)
{
    uint32_t t = r->s[3], s0 = r->s[0];
    r->s[3] = r->s[2];
    r->s[2] = r->s[1];
    r->s[1] = s0;
    t ^= t << 11;
    t ^= t >> 8;
    r->s[0] = t ^ s0 ^ (s0 >> 19);
    return r->s[0];
}
static inline uint32_t hash32(uint32_t x)
{
    x = (x ^ (x >> 16)) * 0x7feb352dU;
    x = (x ^ (x >> 15)) * 0x846ca68bU;
    return x ^ (x >> 16);
}
static void rng_seed(ORng *r, uint32_t seed)
{
    int i;
    for (i = 0; i < 4; ++i) {
        seed = hash32(seed + 0x9e3779b9U * (uint32_t)(i + 1));
        r->s[i] = seed ? seed : 1U;
    }
}
/* u32 * 2^-32; in the FP32 build the u32 is first rounded to float */
static inline real rng_unit(ORng *r) { return (real)rng_next(r) * K(1.0 / 4294967296.0); }
static inline real rng_range(ORng *r, real lo, real hi) { return lo + (hi - lo) * rng_unit(r); }

/* ------------------------------------------------------------------ */
/* terrain, arc_terrain.c:17-71, arc_terrain.h:34-50                    */

static void terrain_flat(OTerrain *t, int nx, int ny, real cell, real height)
{
    int i;
    if (nx > GRID_MAX) nx = GRID_MAX;
    if (ny > GRID_MAX) ny = GRID_MAX;
    t->nx = nx;
    t->ny = ny;
    t->cell = cell;
    t->ox = K(-0.5) * nx * cell;
    t->oy = K(-0.5) * ny * cell;
    t->outside = height;
    for (i = 0; i < nx * ny; ++i)
        t->h[i] = height;
}

static inline real terrain_cell(const OTerrain *t, int ix, int iy)
{
    if (ix < 0 || iy < 0 || ix >= t->nx || iy >= t->ny)
        return t->outside;
    return t->h[ix + iy * t->nx];
}

static inline real terrain_height(const OTerrain *t, real x, real y)
{
    real fx = L1(floor, (x - t->ox) / t->cell), fy = L1(floor, (y - t->oy) / t->cell);
    if (!(fx >= K(0) && fy >= K(0) && fx < t->nx && fy < t->ny))
        return t->outside;
    return t->h[(int)fx + (int)fy * t->nx];
}

static inline real smoothstep3(real t) { return t * t * (K(3.0) - K(2.0) * t); }

static void terrain_noise(OTerrain *t, int spacing, real amp, ORng *r)
{
    real lat[(GRID_MAX / 3 + 3) * (GRID_MAX / 3 + 3)];
    int m = t->nx / spacing + 2, ix, iy, i;
    for (i = 0; i < m * m; ++i)
        lat[i] = rng_range(r, K(-1.0), K(1.0));
    for (iy = 0; iy < t->ny; ++iy)
        for (ix = 0; ix < t->nx; ++ix) {
            real fx = (real)ix / spacing, fy = (real)iy / spacing;
            int gx = (int)fx, gy = (int)fy;
            real tx = smoothstep3(fx - gx), ty = smoothstep3(fy - gy);
            real a = lat[gx + gy * m], b = lat[gx + 1 + gy * m];
            real c = lat[gx + (gy + 1) * m], d = lat[gx + 1 + (gy + 1) * m];
            real v = (a + (b - a) * tx) * (K(1.0) - ty) + (c + (d - c) * tx) * ty;
            t->h[ix + iy * t->nx] += amp * v;
        }
}

static void terrain_generate(OTerrain *t, int n, real cell, real difficulty, ORng *r)
{
    int i;
    real base;
    terrain_flat(t, n, n, cell, K(0.0));
    t->outside = K(-100.0);
    if (difficulty <= K(0.0))
        return;
    terrain_noise(t, 8, K(0.45) * difficulty, r);
    terrain_noise(t, 3, K(0.12) * difficulty, r);
    for (i = 0; i < t->nx * t->ny; ++i) {
        real hh = t->h[i] + K(0.06) * difficulty * rng_range(r, K(-1.0), K(1.0));
        t->h[i] = K(0.03) * L1(floor, hh / K(0.03) + K(0.5));
    }
    base = terrain_height(t, K(0.0), K(0.0));
    for (i = 0; i < t->nx * t->ny; ++i)
        t->h[i] -= base;
}

/* ------------------------------------------------------------------ */
/* bodies and constraint primitives, arc_phys.c:41-371                  */

static inline int iso3(V3 v) { return v.x == v.y && v.y == v.z; }

/* q + 0.5 [w,0] q without renormalising, arc_phys.c:54-61 */
static inline Q4 qturn(Q4 q, V3 w)
{
    real hx = K(0.5) * w.x, hy = K(0.5) * w.y, hz = K(0.5) * w.z;
    return qmk(q.w - hx * q.x - hy * q.y - hz * q.z,
               q.x + hx * q.w + hy * q.z - hz * q.y,
               q.y - hx * q.z + hy * q.w + hz * q.x,
               q.z + hx * q.y - hy * q.x + hz * q.w);
}

/* refresh R and Iw from rq, arc_phys.c:64-82 */
static inline void bsync(OBody *b)
{
    V3 d = b->invI;
    int i, k;
    m_from_q(&b->R, b->rq);
    if (iso3(d)) {
        memset(&b->Iw, 0, sizeof(b->Iw));
        b->Iw.e[0][0] = b->Iw.e[1][1] = b->Iw.e[2][2] = d.x;
        return;
    }
    for (i = 0; i < 3; ++i) {
        V3 rd = vmk(b->R.e[i][0] * d.x, b->R.e[i][1] * d.y, b->R.e[i][2] * d.z);
        for (k = i; k < 3; ++k)
            b->Iw.e[i][k] = b->Iw.e[k][i] =
                rd.x * b->R.e[k][0] + rd.y * b->R.e[k][1] + rd.z * b->R.e[k][2];
    }
}

static inline V3 brot(const OBody *b, V3 v) { return b ? m_mul(&b->R, v) : v; }

static inline real wpos(const OBody *b, V3 r, V3 n)
{
    V3 rn;
    if (!b)
        return K(0.0);
    rn = vcross(r, n);
    return b->invm + vdot(rn, m_mul(&b->Iw, rn));
}
static inline real wrot(const OBody *b, V3 n) { return b ? vdot(n, m_mul(&b->Iw, n)) : K(0.0); }

static inline void push_pos(OBody *b, V3 r, V3 p)
{
    if (!b)
        return;
    b->pos = vmad(b->pos, p, b->invm);
    b->rq = qturn(b->rq, m_mul(&b->Iw, vcross(r, p)));
    bsync(b);
}
static inline void push_rot(OBody *b, V3 p)
{
    if (!b)
        return;
    b->rq = qturn(b->rq, m_mul(&b->Iw, p));
    bsync(b);
}
static inline void push_vel(OBody *b, V3 r, V3 p)
{
    if (!b)
        return;
    b->lin = vmad(b->lin, p, b->invm);
    b->ang = vadd(b->ang, m_mul(&b->Iw, vcross(r, p)));
}

/* arc_phys.c:226-242 */
static real pos_constraint(OBody *b1, V3 r1, OBody *b2, V3 r2, V3 d, real alpha, real h)
{
    real c = vlen(d), den, dl;
    V3 n, p;
    if (c < K(1e-12))
        return K(0.0);
    n = vscl(d, K(1.0) / c);
    den = wpos(b1, r1, n) + wpos(b2, r2, n) + alpha / (h * h);
    if (den <= K(0.0))
        return K(0.0);
    dl = c / den;
    p = vscl(n, dl);
    push_pos(b1, r1, p);
    push_pos(b2, r2, vneg(p));
    return dl;
}

/* Cramer 3x3, arc_phys.c:245-262 */
static int cramer3(real J[3][3], V3 d, V3 *x)
{
    real c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
    real c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
    real c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
    real det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02, inv;
    V3 k0, k1, k2;
    if (fabs((double)det) < (double)K(1e-300))
        return 0;
    inv = K(1.0) / det;
    k0 = vmk(J[0][0], J[1][0], J[2][0]);
    k1 = vmk(J[0][1], J[1][1], J[2][1]);
    k2 = vmk(J[0][2], J[1][2], J[2][2]);
    x->x = vdot(d, vcross(k1, k2)) * inv;
    x->y = vdot(k0, vcross(d, k2)) * inv;
    x->z = vdot(k0, vcross(k1, d)) * inv;
    return 1;
}

/* symmetric 3x3 by adjugate, arc_phys.c:265-280 */
static int sym3(real K3[3][3], V3 d, V3 *x)
{
    real a = K3[0][0], b = K3[0][1], c = K3[0][2], e = K3[1][1], f = K3[1][2], g = K3[2][2];
    real A = e * g - f * f, B = c * f - b * g, C = b * f - c * e;
    real det = a * A + b * B + c * C, inv, D, E, F;
    if (fabs((double)det) < (double)K(1e-300))
        return 0;
    inv = K(1.0) / det;
    D = a * g - c * c;
    E = b * c - a * f;
    F = a * e - b * b;
    x->x = (A * d.x + B * d.y + C * d.z) * inv;
    x->y = (B * d.x + D * d.y + E * d.z) * inv;
    x->z = (C * d.x + E * d.y + F * d.z) * inv;
    return 1;
}

/* point-mass matrix I/m - [r]x Iw [r]x, accumulated, arc_phys.c:284-300 */
static void point_mass(real K3[3][3], const OBody *b, V3 r)
{
    const M33 *a;
    real x = r.x, y = r.y, z = r.z, k01, k02, k12;
    if (!b)
        return;
    a = &b->Iw;
    K3[0][0] += b->invm + a->e[1][1] * z * z - K(2.0) * a->e[1][2] * y * z + a->e[2][2] * y * y;
    K3[1][1] += b->invm + a->e[2][2] * x * x - K(2.0) * a->e[0][2] * x * z + a->e[0][0] * z * z;
    K3[2][2] += b->invm + a->e[0][0] * y * y - K(2.0) * a->e[0][1] * x * y + a->e[1][1] * x * x;
    k01 = a->e[1][2] * x * z + a->e[0][2] * y * z - a->e[0][1] * z * z - a->e[2][2] * x * y;
    k02 = a->e[0][1] * y * z + a->e[1][2] * x * y - a->e[1][1] * x * z - a->e[0][2] * y * y;
    k12 = a->e[0][2] * x * y + a->e[0][1] * x * z - a->e[0][0] * y * z - a->e[1][2] * x * x;
    K3[0][1] += k01; K3[1][0] += k01;
    K3[0][2] += k02; K3[2][0] += k02;
    K3[1][2] += k12; K3[2][1] += k12;
}

/* rigid point-to-point block, arc_phys.c:308-320 */
static void point_block(OBody *b1, V3 r1, OBody *b2, V3 r2, V3 d)
{
    real K3[3][3] = { { 0, 0, 0 }, { 0, 0, 0 }, { 0, 0, 0 } };
    V3 p;
    if (vlen2(d) < K(1e-24))
        return;
    point_mass(K3, b1, r1);
    point_mass(K3, b2, r2);
    if (!sym3(K3, d, &p))
        return;
    push_pos(b1, r1, p);
    push_pos(b2, r2, vneg(p));
}

/* rigid angular block, arc_phys.c:328-346 */
static void rot_block(OBody *b1, OBody *b2, V3 rot)
{
    real K3[3][3] = { { 0, 0, 0 }, { 0, 0, 0 }, { 0, 0, 0 } };
    V3 p;
    int i, k;
    if (vlen2(rot) < K(1e-24))
        return;
    for (i = 0; i < 3; ++i)
        for (k = 0; k < 3; ++k) {
            if (b1)
                K3[i][k] += b1->Iw.e[i][k];
            if (b2)
                K3[i][k] += b2->Iw.e[i][k];
        }
    if (!sym3(K3, rot, &p))
        return;
    push_rot(b1, p);
    push_rot(b2, vneg(p));
}

/* compliant, capped angular constraint, arc_phys.c:353-371 */
static V3 rot_soft(OBody *b1, OBody *b2, V3 rot, real alpha, real h, real cap)
{
    real th = vlen(rot), den, dl;
    V3 n, p;
    if (th < K(1e-12))
        return vmk(K(0), K(0), K(0));
    n = vscl(rot, K(1.0) / th);
    den = wrot(b1, n) + wrot(b2, n) + alpha / (h * h);
    if (den <= K(0.0))
        return vmk(K(0), K(0), K(0));
    dl = th / den;
    if (dl > cap)
        dl = cap;
    p = vscl(n, dl);
    push_rot(b1, p);
    push_rot(b2, vneg(p));
    return p;
}

/* ------------------------------------------------------------------ */
/* joints, arc_phys.c:376-481                                           */

static inline OBody *jparent(OWorld *w, const OJoint *j) { return j->parent >= 0 ? &w->b[j->parent] : 0; }
static inline Q4 bq(const OWorld *w, int i) { return i >= 0 ? w->b[i].rq : qid(); }

static real joint_angle(const OWorld *w, const OJoint *j)
{
    Q4 qp = bq(w, j->parent), qc = bq(w, j->child);
    V3 a = qrot(qp, j->a1), e1 = qrot(qp, j->b1), e2 = qrot(qc, j->b2);
    return (real)atan2((double)vdot(vcross(e1, e2), a), (double)vdot(e1, e2));
}

static real joint_rate(const OWorld *w, const OJoint *j)
{
    V3 a = qrot(bq(w, j->parent), j->a1);
    V3 wp = j->parent >= 0 ? w->b[j->parent].ang : vmk(K(0), K(0), K(0));
    return vdot(vsub(w->b[j->child].ang, wp), a);
}

static inline void joint_sc(const OBody *bp, const OBody *bc, const OJoint *j, V3 a1, real *s, real *c)
{
    V3 e1 = brot(bp, j->b1), e2 = brot(bc, j->b2);
    *s = vdot(vcross(e1, e2), a1);
    *c = vdot(e1, e2);
}

static void joint_positions(OWorld *w, OJoint *j, real h)
{
    OBody *bp = jparent(w, j), *bc = &w->b[j->child];
    V3 a1, r1, r2, p1, p2;
    real phi, s, c;

    if (j->kp > K(0.0)) { /* PD drive, arc_phys.c:422-430 */
        V3 p;
        a1 = brot(bp, j->a1);
        joint_sc(bp, bc, j, a1, &s, &c);
        phi = (real)atan2((double)s, (double)c);
        p = rot_soft(bp, bc, vscl(a1, phi - j->target), K(1.0) / j->kp, h, j->tmax * h * h);
        j->tacc -= vdot(p, a1) / (h * h);
    }
    if (j->lo <= j->hi) { /* limits, arc_phys.c:434-445 */
        a1 = brot(bp, j->a1);
        joint_sc(bp, bc, j, a1, &s, &c);
        if (j->hi - j->lo > PI_R || s * j->lim[0] - c * j->lim[1] < K(0.0)
            || j->lim[3] * c - j->lim[2] * s < K(0.0)) {
            phi = (real)atan2((double)s, (double)c);
            if (phi < j->lo)
                rot_block(bp, bc, vscl(a1, phi - j->lo));
            else if (phi > j->hi)
                rot_block(bp, bc, vscl(a1, phi - j->hi));
        }
    }
    /* axis alignment, arc_phys.c:448-449 */
    a1 = brot(bp, j->a1);
    rot_block(bp, bc, vcross(a1, brot(bc, j->a2)));
    /* attachment, arc_phys.c:452-456 */
    r1 = brot(bp, j->r1);
    r2 = brot(bc, j->r2);
    p1 = bp ? vadd(bp->pos, r1) : r1;
    p2 = vadd(bc->pos, r2);
    point_block(bp, r1, bc, r2, vsub(p2, p1));
}

static void joint_velocities(OWorld *w, OJoint *j, real h)
{
    OBody *bp = jparent(w, j), *bc = &w->b[j->child];
    V3 a, wp;
    real rel, ws, imp, cap;
    if (j->kd <= K(0.0))
        return;
    a = brot(bp, j->a1);
    wp = bp ? bp->ang : vmk(K(0), K(0), K(0));
    rel = vdot(vsub(bc->ang, wp), a);
    ws = wrot(bp, a) + wrot(bc, a);
    if (ws <= K(0.0))
        return;
    imp = -j->kd * rel * h;
    cap = rmin((real)(fabs((double)rel) / (double)ws), j->tmax * h);
    imp = rclamp(imp, -cap, cap);
    if (bp)
        bp->ang = vsub(bp->ang, m_mul(&bp->Iw, vscl(a, imp)));
    bc->ang = vadd(bc->ang, m_mul(&bc->Iw, vscl(a, imp)));
    j->tacc += imp / h;
}

/* ------------------------------------------------------------------ */
/* contacts, arc_phys.c:488-626                                         */

static real contact_solve(OWorld *w, int bi, V3 pt, V3 n, real depth, real h)
{
    OBody *b = &w->b[bi];
    V3 r = vsub(pt, b->pos);
    real ln = pos_constraint(b, r, 0, r, vscl(n, depth), K(0.0), h);
    if (w->nc < NC_MAX) {
        OContact *c = &w->c[w->nc++];
        c->body = bi;
        c->local = m_tmul(&b->R, r);
        c->n = n;
    }
    return ln;
}

static void friction(OWorld *w, OSphere *s, V3 n, real ln, real h)
{
    OBody *b = &w->b[s->body];
    V3 c = vadd(b->pos, brot(b, s->at));
    V3 p = vmad(c, n, -s->rad), e, r;
    real len, wt, lt, frac = K(1.0);
    if (!s->anchored || vdot(n, s->ancn) < K(0.9)) {
        V3 c0 = vadd(b->pos0, qrot(b->rq0, s->at));
        s->anc = vmad(c0, n, -s->rad);
        s->ancn = n;
        s->anchored = 1;
    }
    e = vsub(p, s->anc);
    e = vmad(e, n, -vdot(e, n));
    len = vlen(e);
    if (len < K(1e-12))
        return;
    r = vsub(p, b->pos);
    wt = wpos(b, r, vscl(e, K(1.0) / len));
    lt = len / wt;
    if (lt > w->mus * ln) {
        frac = rmin(w->mud * ln / lt, K(1.0));
        s->anc = vmad(s->anc, e, K(1.0) - frac);
    }
    pos_constraint(b, r, 0, r, vscl(e, -frac), K(0.0), h);
}

static void collide(OWorld *w, OSphere *s, real h)
{
    const OTerrain *t = &w->t;
    OBody *b = &w->b[s->body];
    real rad = s->rad, ln = K(0.0), dmax = K(0.0);
    V3 c = vadd(b->pos, brot(b, s->at)), nmax = vmk(K(0), K(0), K(1));
    int ix0, ix1, iy0, iy1, ix, iy, sx, sy, hit = 0;

    if (!(fabs((double)c.x) < (double)K(1e6) && fabs((double)c.y) < (double)K(1e6))) {
        s->anchored = 0;
        return;
    }
    ix0 = (int)floor((double)((c.x - rad - t->ox) / t->cell));
    ix1 = (int)floor((double)((c.x + rad - t->ox) / t->cell));
    iy0 = (int)floor((double)((c.y - rad - t->oy) / t->cell));
    iy1 = (int)floor((double)((c.y + rad - t->oy) / t->cell));
    for (iy = iy0; iy <= iy1; ++iy)
        for (ix = ix0; ix <= ix1; ++ix) {
            real top = terrain_cell(t, ix, iy);
            real x0 = t->ox + ix * t->cell, x1 = x0 + t->cell;
            real y0 = t->oy + iy * t->cell, y1 = y0 + t->cell;
            real dist, depth;
            V3 q, diff, n;
            if (c.z - rad >= top)
                continue;
            sx = c.x > x1 ? 1 : (c.x < x0 ? -1 : 0);
            sy = c.y > y1 ? 1 : (c.y < y0 ? -1 : 0);
            if (sx && terrain_cell(t, ix + sx, iy) >= top)
                continue;
            if (sy && terrain_cell(t, ix, iy + sy) >= top)
                continue;
            if (sx && sy && terrain_cell(t, ix + sx, iy + sy) >= top)
                continue;
            q = vmk(rclamp(c.x, x0, x1), rclamp(c.y, y0, y1), rmin(c.z, top));
            diff = vsub(c, q);
            dist = vlen(diff);
            if (dist > K(1e-9)) {
                if (dist >= rad)
                    continue;
                n = vscl(diff, K(1.0) / dist);
                depth = rad - dist;
            } else {
                n = vmk(K(0), K(0), K(1));
                depth = top - c.z + rad;
            }
            ln += contact_solve(w, s->body, vmad(c, n, -rad), n, depth, h);
            c = vadd(b->pos, brot(b, s->at));
            if (depth > dmax) {
                dmax = depth;
                nmax = n;
            }
            hit = 1;
        }
    if (hit)
        friction(w, s, nmax, ln, h);
    else
        s->anchored = 0;
    s->touch += hit;
}

static void contact_velocities(OWorld *w)
{
    int i;
    for (i = 0; i < w->nc; ++i) {
        OContact *c = &w->c[i];
        OBody *b = &w->b[c->body];
        V3 r = brot(b, c->local);
        V3 vc = vadd(b->lin, vcross(b->ang, r));
        real vn = vdot(c->n, vc);
        push_vel(b, r, vscl(c->n, -vn / wpos(b, r, c->n)));
    }
}

/* ------------------------------------------------------------------ */
/* stepping, arc_phys.c:636-747                                         */

static V3 gyro(const OBody *b, V3 w0, real h)
{
    real J[3][3], in[3];
    V3 iw = vhad(w0, b->I), f = vscl(vcross(w0, iw), h), dw;
    int k;
    in[0] = b->I.x;
    in[1] = b->I.y;
    in[2] = b->I.z;
    for (k = 0; k < 3; ++k) {
        V3 e = vmk((real)(k == 0), (real)(k == 1), (real)(k == 2)), col;
        col = vadd(vscl(e, in[k]), vscl(vsub(vscl(vcross(w0, e), in[k]), vcross(iw, e)), h));
        J[0][k] = col.x;
        J[1][k] = col.y;
        J[2][k] = col.z;
    }
    if (!cramer3(J, f, &dw))
        return w0;
    return vsub(w0, dw);
}

static void integrate(OWorld *w, real h)
{
    int i;
    for (i = 0; i < w->nb; ++i) {
        OBody *b = &w->b[i];
        b->pos0 = b->pos;
        b->rq0 = b->rq;
        b->lin = vmad(b->lin, w->g, h);
        b->pos = vmad(b->pos, b->lin, h);
        if (!iso3(b->I))
            b->ang = m_mul(&b->R, gyro(b, m_tmul(&b->R, b->ang), h));
        b->rq = qturn(b->rq, vscl(b->ang, h));
        bsync(b);
    }
}

static void velocities(OWorld *w, real h)
{
    int i;
    for (i = 0; i < w->nb; ++i) {
        OBody *b = &w->b[i];
        Q4 dq;
        b->rq = qunit(b->rq);
        dq = qmul(b->rq, qconj(b->rq0));
        b->lin = vscl(vsub(b->pos, b->pos0), K(1.0) / h);
        b->ang = vscl(vmk(dq.x, dq.y, dq.z), K(2.0) / h);
        if (dq.w < K(0.0))
            b->ang = vneg(b->ang);
    }
}

static inline int finite_r(real v) { return v == v && v < K(1e300) && v > -K(1e300); }

static void stability(OWorld *w)
{
    int i;
    for (i = 0; i < w->nb; ++i) {
        OBody *b = &w->b[i];
        real v2 = vlen2(b->lin), w2 = vlen2(b->ang);
        if (!finite_r(v2) || !finite_r(w2) || !(vlen2(b->pos) < K(1e6) * K(1e6))
            || v2 > K(60.0) * K(60.0) || w2 > K(400.0) * K(400.0))
            w->unstable = 1;
    }
}

static void world_step(OWorld *w, real dt, int substeps)
{
    real h = dt / substeps;
    int s, i;
    for (i = 0; i < w->nb; ++i)
        bsync(&w->b[i]);
    for (i = 0; i < w->ns; ++i)
        w->s[i].touch = 0;
    for (i = 0; i < w->nj; ++i) {
        OJoint *j = &w->j[i];
        j->tacc = K(0.0);
        j->lim[0] = L1(cos, j->lo);
        j->lim[1] = L1(sin, j->lo);
        j->lim[2] = L1(cos, j->hi);
        j->lim[3] = L1(sin, j->hi);
    }
    for (s = 0; s < substeps; ++s) {
        integrate(w, h);
        w->nc = 0;
        for (i = 0; i < w->nj; ++i)
            joint_positions(w, &w->j[i], h);
        for (i = 0; i < w->ns; ++i)
            collide(w, &w->s[i], h);
        velocities(w, h);
        for (i = 0; i < w->nj; ++i)
            joint_velocities(w, &w->j[i], h);
        contact_velocities(w);
    }
    for (i = 0; i < w->nj; ++i) {
        OJoint *j = &w->j[i];
        j->angle = joint_angle(w, j);
        j->rate = joint_rate(w, j);
        j->torque = j->tacc / substeps;
    }
    w->time += dt;
    stability(w);
}

/* forward kinematics along the joint order, arc_phys.c:749-782 */
static void world_pose(OWorld *w, V3 pos, Q4 q, const real *angles)
{
    int i;
    OBody *root = &w->b[0];
    root->pos = pos;
    root->rq = qunit(q);
    for (i = 0; i < w->nj; ++i) {
        OJoint *j = &w->j[i];
        OBody *c = &w->b[j->child];
        V3 xp = j->parent >= 0 ? w->b[j->parent].pos : vmk(K(0), K(0), K(0));
        Q4 qp = bq(w, j->parent);
        V3 anchor = vadd(xp, qrot(qp, j->r1));
        c->rq = qunit(qmul(qp, qmul(qaxis(j->a1, angles[i]), j->qrel0)));
        c->pos = vsub(anchor, qrot(c->rq, j->r2));
        j->target = angles[i];
        j->angle = angles[i];
        j->rate = K(0.0);
        j->torque = K(0.0);
    }
    for (i = 0; i < w->nb; ++i) {
        OBody *b = &w->b[i];
        b->pos0 = b->pos;
        b->rq0 = b->rq;
        b->lin = vmk(K(0), K(0), K(0));
        b->ang = vmk(K(0), K(0), K(0));
        bsync(b);
    }
    for (i = 0; i < w->ns; ++i) {
        w->s[i].touch = 0;
        w->s[i].anchored = 0;
    }
    w->nc = 0;
    w->unstable = 0;
}

static real world_energy(const OWorld *w)
{
    real e = K(0.0);
    int i;
    for (i = 0; i < w->nb; ++i) {
        const OBody *b = &w->b[i];
        V3 wl = qirot(b->rq, b->ang);
        e += K(0.5) * b->mass * vlen2(b->lin);
        e += K(0.5) * vdot(wl, vhad(wl, b->I));
        e -= b->mass * vdot(w->g, b->pos);
    }
    return e;
}

/* ------------------------------------------------------------------ */
/* builders, arc_phys.c:32-143, 801-821                                 */

static void world_init(OWorld *w)
{
    memset(w, 0, sizeof(*w));
    w->g = vmk(K(0.0), K(0.0), K(-9.81));
    w->mus = K(1.0);
    w->mud = K(0.8);
    terrain_flat(&w->t, 1, 1, K(1.0), K(0.0));
}

static int add_body(OWorld *w, real mass, V3 I, V3 pos, Q4 q)
{
    OBody *b;
    if (w->nb >= NB_MAX)
        return -1;
    b = &w->b[w->nb];
    memset(b, 0, sizeof(*b));
    b->pos = b->pos0 = pos;
    b->rq = b->rq0 = qunit(q);
    b->mass = mass;
    b->invm = K(1.0) / mass;
    b->I = I;
    b->invI = vmk(K(1.0) / I.x, K(1.0) / I.y, K(1.0) / I.z);
    bsync(b);
    return w->nb++;
}

static int add_sphere(OWorld *w, int body, V3 at, real rad)
{
    OSphere *s;
    if (w->ns >= NS_MAX)
        return -1;
    s = &w->s[w->ns];
    memset(s, 0, sizeof(*s));
    s->body = body;
    s->at = at;
    s->rad = rad;
    return w->ns++;
}

static int add_hinge(OWorld *w, int parent, int child, V3 anchor, V3 axis_w)
{
    OJoint *j;
    Q4 qp = qid(), qc;
    V3 xp = vmk(K(0), K(0), K(0)), axis, ref;
    if (w->nj >= NJ_MAX || child <= parent)
        return -1;
    if (parent >= 0) {
        xp = w->b[parent].pos;
        qp = w->b[parent].rq;
    }
    qc = w->b[child].rq;
    axis = vunit(axis_w);
    ref = vperp(axis);
    j = &w->j[w->nj];
    memset(j, 0, sizeof(*j));
    j->parent = parent;
    j->child = child;
    j->r1 = qirot(qp, vsub(anchor, xp));
    j->r2 = qirot(qc, vsub(anchor, w->b[child].pos));
    j->a1 = qirot(qp, axis);
    j->a2 = qirot(qc, axis);
    j->b1 = qirot(qp, ref);
    j->b2 = qirot(qc, ref);
    j->qrel0 = qmul(qconj(qp), qc);
    j->lo = K(1.0);
    j->hi = K(-1.0);
    return w->nj++;
}

static V3 inertia_box(real m, V3 he)
{
    real k = m / K(3.0);
    return vmk(k * (he.y * he.y + he.z * he.z), k * (he.x * he.x + he.z * he.z),
               k * (he.x * he.x + he.y * he.y));
}
static V3 inertia_capsule_z(real m, real r, real half)
{
    real len = K(2.0) * (half + r);
    real ixx = m * (K(3.0) * r * r + len * len) / K(12.0);
    return vmk(ixx, ixx, K(0.5) * m * r * r);
}
static V3 inertia_sphere(real m, real r)
{
    real i = K(0.4) * m * r * r;
    return vmk(i, i, i);
}

/* ------------------------------------------------------------------ */
/* the quadruped, arc_robot.c:51-70, 98-144, 180-229 (no visuals)       */

typedef struct {
    int foot[LEGS];
    int n_torso_spheres;
    real def[NJR], lo[NJR], hi[NJR];
    real stand, mass, footprint;
} ORobotInfo;

static void robot_leg(OWorld *w, ORobotInfo *info, int leg, real sx, real sy)
{
    const real R0 = K(0.2), ELEV = K(0.785), SPLAY = K(0.785);
    const real LC = K(0.06), LF = K(0.50), LT = K(0.95);
    real psi = sx > K(0.0) ? sy * SPLAY : sy * (PI_R - SPLAY);
    V3 d = vmk(L1(cos, psi), L1(sin, psi), K(0.0));
    V3 up = vmk(K(0), K(0), K(1));
    V3 lift = vcross(d, up);
    Q4 along = qaxis(vunit(vcross(up, d)), K(0.5) * PI_R);
    /* BODY_R * cos(HIP_ELEVATION) is a double product in both builds */
    V3 hip = vadd(vscl(d, (real)((double)R0 * cos((double)ELEV))),
                  vmk(K(0), K(0), (real)((double)R0 * sin((double)ELEV))));
    V3 p1 = vmad(hip, d, LC), p2 = vmad(p1, d, LF);
    real fl = K(0.5) * LF, tl = K(0.5) * LT;
    int bc, bf, bt, j;

    bc = add_body(w, K(0.4), vmk(K(0.02), K(0.02), K(0.02)), vmad(hip, d, K(0.5) * LC), qid());
    bf = add_body(w, K(1.5), inertia_capsule_z(K(1.5), K(0.06), fl), vmad(p1, d, fl), along);
    bt = add_body(w, K(1.0), inertia_capsule_z(K(1.0), K(0.05), tl), vmad(p2, d, tl), along);

    j = add_hinge(w, 0, bc, hip, vmk(K(0.0), K(0.0), -sy));
    w->j[j].lo = K(-0.8); w->j[j].hi = K(0.8);
    j = add_hinge(w, bc, bf, p1, lift);
    w->j[j].lo = K(-0.4); w->j[j].hi = K(1.5);
    j = add_hinge(w, bf, bt, p2, lift);
    w->j[j].lo = K(-2.8); w->j[j].hi = K(-0.5);
    for (j = w->nj - 3; j < w->nj; ++j) {
        w->j[j].kp = K(250.0);
        w->j[j].kd = K(6.0);
        w->j[j].tmax = K(80.0);
    }
    info->def[3 * leg + 0] = K(0.0);
    info->def[3 * leg + 1] = K(0.6);
    info->def[3 * leg + 2] = K(-1.95);

    add_sphere(w, bf, vmk(K(0), K(0), K(0)), K(0.06));
    add_sphere(w, bt, vmk(K(0), K(0), -tl), K(0.07));
    add_sphere(w, bt, vmk(K(0), K(0), K(-0.3) * tl), K(0.055));
    add_sphere(w, bt, vmk(K(0), K(0), K(0.4) * tl), K(0.035));
    info->foot[leg] = add_sphere(w, bt, vmk(K(0), K(0), tl), K(0.045));
}

static void robot_build(OWorld *w, ORobotInfo *info)
{
    static const double SX[LEGS] = { 1.0, 1.0, -1.0, -1.0 };
    static const double SY[LEGS] = { 1.0, -1.0, 1.0, -1.0 };
    real lowest = K(1e30);
    int leg, i;

    add_body(w, K(8.0), inertia_sphere(K(8.0), K(0.2)), vmk(K(0), K(0), K(0)), qid());
    add_sphere(w, 0, vmk(K(0), K(0), K(0)), K(0.2));
    info->n_torso_spheres = w->ns;
    for (leg = 0; leg < LEGS; ++leg)
        robot_leg(w, info, leg, (real)SX[leg], (real)SY[leg]);
    for (i = 0; i < NJR; ++i) {
        info->lo[i] = w->j[i].lo;
        info->hi[i] = w->j[i].hi;
    }
    info->mass = K(0.0);
    for (i = 0; i < w->nb; ++i)
        info->mass += w->b[i].mass;
    world_pose(w, vmk(K(0), K(0), K(0)), qid(), info->def);
    info->footprint = K(0.0);
    for (leg = 0; leg < LEGS; ++leg) {
        const OSphere *s = &w->s[info->foot[leg]];
        const OBody *b = &w->b[s->body];
        V3 c = vadd(b->pos, qrot(b->rq, s->at));
        lowest = rmin(lowest, c.z - s->rad);
        /* sqrt(...) / 4 and the accumulation are double in both builds */
        info->footprint = (real)((double)info->footprint + sqrt((double)(c.x * c.x + c.y * c.y)) / LEGS);
    }
    info->stand = -lowest;
    world_pose(w, vmk(K(0), K(0), info->stand), qid(), info->def);
}

/* ------------------------------------------------------------------ */
/* batched sim, arc_sim.c:26-215                                        */

typedef struct {
    int n, substeps, grid;
    real dt, cell;
    double half_extent;  /* task fall rule uses the caller's double cell */
    ORobotInfo info;
    int channel[NS_MAX];
    OWorld *tmpl;
    OWorld *worlds;
} OSim;

/* reset one world, arc_sim.c:88-110 */
static void sim_reset_world(const OSim *s, OWorld *w, uint32_t seed, real difficulty)
{
    ORng r;
    real ang[NJR], ground = K(-1e30), yaw, x, y, reach;
    int i;
    rng_seed(&r, seed);
    memcpy(w, s->tmpl, offsetof(OWorld, t));
    terrain_generate(&w->t, s->grid, s->cell, difficulty, &r);
    for (i = 0; i < NJR; ++i)
        ang[i] = rclamp(s->info.def[i] + rng_range(&r, K(-0.1), K(0.1)), s->info.lo[i], s->info.hi[i]);
    yaw = rng_range(&r, -PI_R, PI_R);
    reach = s->info.footprint + K(0.1);
    for (x = -reach; x <= reach; x += K(0.5) * s->cell)
        for (y = -reach; y <= reach; y += K(0.5) * s->cell)
            ground = rmax(ground, terrain_height(&w->t, x, y));
    world_pose(w, vmk(K(0.0), K(0.0), ground + s->info.stand + K(0.03)),
               qaxis(vmk(K(0.0), K(0.0), K(1.0)), yaw), ang);
}

/* ================================================================== */
/* exported API (double at the boundary)                              */

#define EXPORT __attribute__((visibility("default")))

EXPORT int orc_real_bytes(void) { return (int)sizeof(real); }
EXPORT int orc_world_bytes(void) { return (int)sizeof(OWorld); }

/* --- RNG ------------------------------------------------------------ */
EXPORT void orc_rng_stream(uint32_t seed, int n, uint32_t *out)
{
    ORng r;
    int i;
    rng_seed(&r, seed);
    for (i = 0; i < n; ++i)
        out[i] = rng_next(&r);
}
EXPORT uint32_t orc_hash32(uint32_t x) { return hash32(x); }
EXPORT void orc_rng_uniform(uint32_t seed, int n, double *out)
{
    ORng r;
    int i;
    rng_seed(&r, seed);
    for (i = 0; i < n; ++i)
        out[i] = (double)rng_unit(&r);
}

/* --- single worlds (known-answer tests of arc_phys) ------------------ */
EXPORT OWorld *orc_world_new(void)
{
    OWorld *w = (OWorld *)malloc(sizeof(OWorld));
    if (w)
        world_init(w);
    return w;
}
EXPORT void orc_world_free(OWorld *w) { free(w); }
EXPORT void orc_world_set_gravity(OWorld *w, double x, double y, double z) { w->g = vmk((real)x, (real)y, (real)z); }
EXPORT void orc_world_set_mu(OWorld *w, double ms, double md) { w->mus = (real)ms; w->mud = (real)md; }
EXPORT void orc_terrain_flat(OWorld *w, int nx, int ny, double cell, double height)
{
    terrain_flat(&w->t, nx, ny, (real)cell, (real)height);
}
EXPORT void orc_terrain_generate(OWorld *w, int n, double cell, double difficulty, uint32_t seed)
{
    ORng r;
    rng_seed(&r, seed);
    terrain_generate(&w->t, n, (real)cell, (real)difficulty, &r);
}
EXPORT void orc_terrain_set(OWorld *w, int i, double h) { w->t.h[i] = (real)h; }
EXPORT double orc_terrain_height(const OWorld *w, double x, double y) { return (double)terrain_height(&w->t, (real)x, (real)y); }
EXPORT void orc_terrain_info(const OWorld *w, double *out)
{
    out[0] = w->t.nx; out[1] = w->t.ny; out[2] = (double)w->t.cell;
    out[3] = (double)w->t.ox; out[4] = (double)w->t.oy; out[5] = (double)w->t.outside;
}
EXPORT int orc_add_body(OWorld *w, double mass, const double *I3, const double *p3, const double *q4)
{
    return add_body(w, (real)mass, vmk((real)I3[0], (real)I3[1], (real)I3[2]),
                    vmk((real)p3[0], (real)p3[1], (real)p3[2]),
                    qmk((real)q4[0], (real)q4[1], (real)q4[2], (real)q4[3]));
}
EXPORT int orc_add_sphere(OWorld *w, int body, const double *p3, double rad)
{
    return add_sphere(w, body, vmk((real)p3[0], (real)p3[1], (real)p3[2]), (real)rad);
}
EXPORT int orc_add_hinge(OWorld *w, int parent, int child, const double *anchor3, const double *axis3)
{
    return add_hinge(w, parent, child, vmk((real)anchor3[0], (real)anchor3[1], (real)anchor3[2]),
                     vmk((real)axis3[0], (real)axis3[1], (real)axis3[2]));
}
/* lo, hi, kp, kd, max_torque, target */
EXPORT void orc_joint_set(OWorld *w, int j, const double *p6)
{
    OJoint *jt = &w->j[j];
    jt->lo = (real)p6[0]; jt->hi = (real)p6[1]; jt->kp = (real)p6[2];
    jt->kd = (real)p6[3]; jt->tmax = (real)p6[4]; jt->target = (real)p6[5];
}
/* angle, rate, torque */
EXPORT void orc_joint_out(const OWorld *w, int j, double *out3)
{
    out3[0] = (double)w->j[j].angle; out3[1] = (double)w->j[j].rate; out3[2] = (double)w->j[j].torque;
}
EXPORT double orc_joint_angle(const OWorld *w, int j) { return (double)joint_angle(w, &w->j[j]); }
EXPORT double orc_joint_rate(const OWorld *w, int j) { return (double)joint_rate(w, &w->j[j]); }
/* x(3) q(4) v(3) w(3) */
EXPORT void orc_body_get(const OWorld *w, int i, double *out13)
{
    const OBody *b = &w->b[i];
    out13[0] = (double)b->pos.x; out13[1] = (double)b->pos.y; out13[2] = (double)b->pos.z;
    out13[3] = (double)b->rq.w; out13[4] = (double)b->rq.x; out13[5] = (double)b->rq.y; out13[6] = (double)b->rq.z;
    out13[7] = (double)b->lin.x; out13[8] = (double)b->lin.y; out13[9] = (double)b->lin.z;
    out13[10] = (double)b->ang.x; out13[11] = (double)b->ang.y; out13[12] = (double)b->ang.z;
}
EXPORT void orc_body_set(OWorld *w, int i, const double *in13)
{
    OBody *b = &w->b[i];
    b->pos = vmk((real)in13[0], (real)in13[1], (real)in13[2]);
    b->rq = qmk((real)in13[3], (real)in13[4], (real)in13[5], (real)in13[6]);
    b->lin = vmk((real)in13[7], (real)in13[8], (real)in13[9]);
    b->ang = vmk((real)in13[10], (real)in13[11], (real)in13[12]);
}
EXPORT void orc_body_inertia(const OWorld *w, int i, double *out3)
{
    out3[0] = (double)w->b[i].I.x; out3[1] = (double)w->b[i].I.y; out3[2] = (double)w->b[i].I.z;
}
EXPORT void orc_world_step(OWorld *w, double dt, int substeps, int steps)
{
    int k;
    for (k = 0; k < steps; ++k)
        world_step(w, (real)dt, substeps);
}
EXPORT void orc_world_set_pose(OWorld *w, const double *p3, const double *q4, const double *angles)
{
    real a[NJ_MAX];
    int i;
    for (i = 0; i < w->nj; ++i)
        a[i] = (real)angles[i];
    world_pose(w, vmk((real)p3[0], (real)p3[1], (real)p3[2]),
               qmk((real)q4[0], (real)q4[1], (real)q4[2], (real)q4[3]), a);
}
EXPORT double orc_world_energy(const OWorld *w) { return (double)world_energy(w); }
EXPORT int orc_world_unstable(const OWorld *w) { return w->unstable; }
EXPORT double orc_world_time(const OWorld *w) { return (double)w->time; }
EXPORT int orc_world_counts(const OWorld *w, int *out4)
{
    out4[0] = w->nb; out4[1] = w->nj; out4[2] = w->ns; out4[3] = w->nc;
    return w->unstable;
}
/* touch, anchored, anchor(3), anchor_n(3) */
EXPORT void orc_sphere_get(const OWorld *w, int i, double *out8)
{
    const OSphere *s = &w->s[i];
    out8[0] = s->touch; out8[1] = s->anchored;
    out8[2] = (double)s->anc.x; out8[3] = (double)s->anc.y; out8[4] = (double)s->anc.z;
    out8[5] = (double)s->ancn.x; out8[6] = (double)s->ancn.y; out8[7] = (double)s->ancn.z;
}
EXPORT void orc_inertia(int kind, double mass, double a, double b, double c, double *out3)
{
    V3 r = kind == 0 ? inertia_sphere((real)mass, (real)a)
         : kind == 1 ? inertia_capsule_z((real)mass, (real)a, (real)b)
                     : inertia_box((real)mass, vmk((real)a, (real)b, (real)c));
    out3[0] = (double)r.x; out3[1] = (double)r.y; out3[2] = (double)r.z;
}

/* --- batched sim (arc_sim.h) ----------------------------------------- */
EXPORT OSim *orc_sim_create(int n, double dt, int substeps, int grid, double cell)
{
    OSim *s = (OSim *)calloc(1, sizeof(OSim));
    int i;
    if (!s)
        return 0;
    s->n = n;
    s->dt = (real)dt;
    s->substeps = substeps;
    s->grid = grid < GRID_MAX ? grid : GRID_MAX;
    s->cell = (real)cell;
    s->half_extent = 0.5 * s->grid * cell;
    s->tmpl = (OWorld *)malloc(sizeof(OWorld));
    s->worlds = (OWorld *)malloc(sizeof(OWorld) * (size_t)(n > 0 ? n : 1));
    if (!s->tmpl || !s->worlds) {
        free(s->tmpl); free(s->worlds); free(s);
        return 0;
    }
    world_init(s->tmpl);
    robot_build(s->tmpl, &s->info);
    terrain_flat(&s->tmpl->t, s->grid, s->grid, s->cell, K(0.0));
    for (i = 0; i < s->tmpl->ns; ++i)
        s->channel[i] = i < s->info.n_torso_spheres ? LEGS : LEGS + 1;
    for (i = 0; i < LEGS; ++i)
        s->channel[s->info.foot[i]] = i;
    for (i = 0; i < n; ++i)
        memcpy(&s->worlds[i], s->tmpl, sizeof(OWorld));
    return s;
}
EXPORT void orc_sim_destroy(OSim *s)
{
    if (!s)
        return;
    free(s->tmpl);
    free(s->worlds);
    free(s);
}
EXPORT int orc_sim_num_worlds(const OSim *s) { return s->n; }
/* default(12) lo(12) hi(12) stand mass footprint */
EXPORT void orc_sim_robot_info(const OSim *s, double *def, double *lo, double *hi, double *misc3)
{
    int i;
    for (i = 0; i < NJR; ++i) {
        if (def) def[i] = (double)s->info.def[i];
        if (lo) lo[i] = (double)s->info.lo[i];
        if (hi) hi[i] = (double)s->info.hi[i];
    }
    if (misc3) {
        misc3[0] = (double)s->info.stand;
        misc3[1] = (double)s->info.mass;
        misc3[2] = (double)s->info.footprint;
    }
}
EXPORT OWorld *orc_sim_world(OSim *s, int i) { return i < 0 ? s->tmpl : &s->worlds[i]; }
EXPORT void orc_sim_reset(OSim *s, const int *ids, int n_ids, const uint32_t *seeds, const double *difficulty)
{
    int k;
#pragma omp parallel for schedule(static)
    for (k = 0; k < n_ids; ++k)
        sim_reset_world(s, &s->worlds[ids[k]], seeds[k], (real)difficulty[k]);
}
EXPORT void orc_sim_step(OSim *s, const double *targets, int threads)
{
    int i;
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
    for (i = 0; i < s->n; ++i) {
        OWorld *w = &s->worlds[i];
        int j;
        for (j = 0; j < NJR; ++j)
            w->j[j].target = rclamp((real)targets[i * NJR + j], w->j[j].lo, w->j[j].hi);
        if (!w->unstable)
            world_step(w, s->dt, s->substeps);
    }
}
EXPORT void orc_sim_observe(const OSim *s, double *root, double *rot, double *joints, double *contacts, int *flags)
{
    int i;
    for (i = 0; i < s->n; ++i) {
        const OWorld *w = &s->worlds[i];
        const OBody *b = &w->b[0];
        int j;
        if (root) {
            double *r = root + i * 13;
            r[0] = (double)b->pos.x; r[1] = (double)b->pos.y; r[2] = (double)b->pos.z;
            r[3] = (double)b->rq.w; r[4] = (double)b->rq.x; r[5] = (double)b->rq.y; r[6] = (double)b->rq.z;
            r[7] = (double)b->lin.x; r[8] = (double)b->lin.y; r[9] = (double)b->lin.z;
            r[10] = (double)b->ang.x; r[11] = (double)b->ang.y; r[12] = (double)b->ang.z;
        }
        if (rot) {
            M33 m;
            int a, c;
            m_from_q(&m, b->rq);
            for (a = 0; a < 3; ++a)
                for (c = 0; c < 3; ++c)
                    rot[i * 9 + a * 3 + c] = (double)m.e[a][c];
        }
        if (joints) {
            double *o = joints + i * 36;
            for (j = 0; j < NJR; ++j) {
                o[j] = (double)w->j[j].angle;
                o[NJR + j] = (double)w->j[j].rate;
                o[2 * NJR + j] = (double)w->j[j].torque;
            }
        }
        if (contacts) {
            real o[LEGS + 2], inv = K(1.0) / s->substeps;
            for (j = 0; j < LEGS + 2; ++j)
                o[j] = K(0.0);
            for (j = 0; j < w->ns; ++j) {
                real *c = &o[s->channel[j]];
                *c = rmax(*c, w->s[j].touch * inv);
            }
            for (j = 0; j < LEGS + 2; ++j)
                contacts[i * 6 + j] = (double)o[j];
        }
        if (flags)
            flags[i] = w->unstable;
    }
}
EXPORT void orc_sim_heights(const OSim *s, const double *offsets, int k, double *out)
{
    int i;
    for (i = 0; i < s->n; ++i) {
        const OWorld *w = &s->worlds[i];
        const OBody *b = &w->b[0];
        V3 fwd = qrot(b->rq, vmk(K(1.0), K(0.0), K(0.0)));
        real yaw = (real)atan2((double)fwd.y, (double)fwd.x), c = L1(cos, yaw), sn = L1(sin, yaw);
        int j;
        for (j = 0; j < k; ++j) {
            real f = (real)offsets[2 * j], l = (real)offsets[2 * j + 1];
            out[i * k + j] = (double)terrain_height(&w->t, b->pos.x + c * f - sn * l, b->pos.y + sn * f + c * l);
        }
    }
}
EXPORT void orc_sim_terrain_at(const OSim *s, const int *ids, const double *xy, int m, double *out)
{
    int i;
    for (i = 0; i < m; ++i)
        out[i] = (double)terrain_height(&s->worlds[ids[i]].t, (real)xy[2 * i], (real)xy[2 * i + 1]);
}
/*
 * Per-world dynamic state in a flat double layout shared with the device
 * engine's dump: 13 bodies x (x3 q4 v3 w3), 12 joints x (target angle rate
 * torque), 21 spheres x (touch anchored anchor3 anchor_n3), unstable, time.
 */
EXPORT int orc_state_size(void) { return 13 * 13 + 12 * 4 + 21 * 8 + 2; }
EXPORT void orc_sim_state(const OSim *s, int i, double *out)
{
    const OWorld *w = i < 0 ? s->tmpl : &s->worlds[i];
    int k, o = 0;
    for (k = 0; k < 13; ++k) {
        orc_body_get(w, k, out + o);
        o += 13;
    }
    for (k = 0; k < 12; ++k) {
        out[o++] = (double)w->j[k].target; out[o++] = (double)w->j[k].angle;
        out[o++] = (double)w->j[k].rate; out[o++] = (double)w->j[k].torque;
    }
    for (k = 0; k < 21; ++k) {
        orc_sphere_get(w, k, out + o);
        o += 8;
    }
    out[o++] = w->unstable;
    out[o++] = (double)w->time;
}
/* nx*ny heights then outside */
EXPORT void orc_sim_terrain(const OSim *s, int i, double *out)
{
    const OTerrain *t = &s->worlds[i].t;
    int k;
    for (k = 0; k < t->nx * t->ny; ++k)
        out[k] = (double)t->h[k];
    out[t->nx * t->ny] = (double)t->outside;
}
EXPORT int orc_sim_num_contacts(const OSim *s, int i) { return s->worlds[i].nc; }

/*
 * Fall rules of the task, arc/task.py:180-203, in the precedence
 * bad > void > flip > beached.  torso_steps is the per-world counter of
 * consecutive steps with torso contact; cause: 0 none, 1 unstable/non-finite,
 * 2 void, 3 flip, 4 beached.  Comparisons are made in double, as the task
 * makes them on the float64 observation.
 */
EXPORT void orc_sim_falls(const OSim *s, int *torso_steps, int *cause, int beached_steps,
                          double flip_up_z, double void_z)
{
    int i;
    for (i = 0; i < s->n; ++i) {
        const OWorld *w = &s->worlds[i];
        const OBody *b = &w->b[0];
        M33 m;
        real touch_frac = K(0.0), inv = K(1.0) / s->substeps;
        int j, bad, vd, flip, beach;
        double r[13];
        orc_body_get(w, 0, r);
        bad = w->unstable != 0;
        for (j = 0; j < 13; ++j)
            bad |= !isfinite(r[j]);
        for (j = 0; j < w->ns; ++j)
            if (s->channel[j] == LEGS)
                touch_frac = rmax(touch_frac, w->s[j].touch * inv);
        torso_steps[i] = touch_frac > K(0.0) ? torso_steps[i] + 1 : 0;
        m_from_q(&m, b->rq);
        vd = fabs(r[0]) > s->half_extent || fabs(r[1]) > s->half_extent || r[2] < void_z;
        flip = (double)m.e[2][2] < flip_up_z;
        beach = torso_steps[i] > beached_steps;
        cause[i] = bad ? 1 : vd ? 2 : flip ? 3 : beach ? 4 : 0;
    }
}
