/*
 * Host-side initial conditions: dSFMT-19937 and the Plummer sampler.
 *
 * dSFMT is written from the published algorithm (Saito & Matsumoto 2009);
 * its parameter set for exponent 19937 (shift/mask/fix/parity constants
 * below) is the published one and must match for the streams to agree.
 * The stream is pinned by the reference's known-answer file
 * dSFMT/dSFMT.19937.out.txt.
 */
#include "../../include/nbmodels.h"

#include <math.h>
#include <string.h>

#define NBM_N64 (2 * NBM_DSFMT_N)
#define NBM_POS1 117
#define NBM_SL1 19
#define NBM_SR 12
#define NBM_MSK1 0x000ffafffffffb3fULL
#define NBM_MSK2 0x000ffdfffc90fffdULL
#define NBM_FIX1 0x90014964b32f4329ULL
#define NBM_FIX2 0x3b8d12ac548a7c7aULL
#define NBM_PCV1 0x3d84e1ac0dc82880ULL
#define NBM_PCV2 0x0000000000000001ULL
#define NBM_LOW_MASK 0x000FFFFFFFFFFFFFULL
#define NBM_HIGH_CONST 0x3FF0000000000000ULL

/* One step of the recursion on 128-bit words a (= r), b, with the lung word. */
static inline void nbm_recursion(uint64_t* r, const uint64_t* b, uint64_t* lung)
{
    uint64_t t0 = r[0], t1 = r[1];
    uint64_t L0 = lung[0], L1 = lung[1];
    lung[0] = (t0 << NBM_SL1) ^ (L1 >> 32) ^ (L1 << 32) ^ b[0];
    lung[1] = (t1 << NBM_SL1) ^ (L0 >> 32) ^ (L0 << 32) ^ b[1];
    r[0] = (lung[0] >> NBM_SR) ^ (lung[0] & NBM_MSK1) ^ t0;
    r[1] = (lung[1] >> NBM_SR) ^ (lung[1] & NBM_MSK2) ^ t1;
}

static void nbm_gen_all(nbm_dsfmt* st)
{
    uint64_t* s = st->status;
    uint64_t lung[2] = { s[2 * NBM_DSFMT_N], s[2 * NBM_DSFMT_N + 1] };
    int i;
    for (i = 0; i < NBM_DSFMT_N - NBM_POS1; ++i)
        nbm_recursion(&s[2 * i], &s[2 * (i + NBM_POS1)], lung);
    for (; i < NBM_DSFMT_N; ++i)
        nbm_recursion(&s[2 * i], &s[2 * (i + NBM_POS1 - NBM_DSFMT_N)], lung);
    s[2 * NBM_DSFMT_N] = lung[0];
    s[2 * NBM_DSFMT_N + 1] = lung[1];
}

void nbm_dsfmt_init(nbm_dsfmt* st, uint32_t seed)
{
    uint32_t w[4 * (NBM_DSFMT_N + 1)];
    w[0] = seed;
    for (uint32_t i = 1; i < 4u * (NBM_DSFMT_N + 1); ++i)
        w[i] = 1812433253u * (w[i - 1] ^ (w[i - 1] >> 30)) + i;
    /* little-endian: 32-bit word 2k is the low half of 64-bit word k */
    for (int k = 0; k < 2 * (NBM_DSFMT_N + 1); ++k)
        st->status[k] = (uint64_t) w[2 * k] | ((uint64_t) w[2 * k + 1] << 32);
    for (int k = 0; k < NBM_N64; ++k)
        st->status[k] = (st->status[k] & NBM_LOW_MASK) | NBM_HIGH_CONST;
    /* period certification */
    uint64_t t0 = st->status[2 * NBM_DSFMT_N] ^ NBM_FIX1;
    uint64_t t1 = st->status[2 * NBM_DSFMT_N + 1] ^ NBM_FIX2;
    uint64_t inner = (t0 & NBM_PCV1) ^ (t1 & NBM_PCV2);
    for (int sh = 32; sh > 0; sh >>= 1) inner ^= inner >> sh;
    if ((inner & 1u) == 0) st->status[2 * NBM_DSFMT_N + 1] ^= 1u;
    st->idx = NBM_N64;
}

double nbm_dsfmt_close1_open2(nbm_dsfmt* st)
{
    if (st->idx >= NBM_N64) {
        nbm_gen_all(st);
        st->idx = 0;
    }
    double r;
    memcpy(&r, &st->status[st->idx++], sizeof(r));
    return r;
}

double nbm_dsfmt_close_open(nbm_dsfmt* st) { return nbm_dsfmt_close1_open2(st) - 1.0; }

double nbm_dsfmt_open_close(nbm_dsfmt* st) { return 2.0 - nbm_dsfmt_close1_open2(st); }

double nbm_dsfmt_open_open(nbm_dsfmt* st)
{
    double d = nbm_dsfmt_close1_open2(st);
    uint64_t u;
    memcpy(&u, &d, sizeof(u));
    u |= 1u;
    memcpy(&d, &u, sizeof(d));
    return d - 1.0;
}

void nbm_dsfmt_fill(nbm_dsfmt* st, int kind, int n, double* out)
{
    for (int i = 0; i < n; ++i) {
        switch (kind) {
        case 0: out[i] = nbm_dsfmt_close1_open2(st); break;
        case 1: out[i] = nbm_dsfmt_close_open(st); break;
        case 2: out[i] = nbm_dsfmt_open_close(st); break;
        default: out[i] = nbm_dsfmt_open_open(st); break;
        }
    }
}

/* mwXrandom (milkyway/include/milkyway_util.h:209) */
static inline double nbm_xrandom(nbm_dsfmt* st, double xl, double xh)
{
    return xl + (xh - xl) * nbm_dsfmt_open_open(st);
}

/* pickShell (nbody/src/nbody_plummer.c:29-45) with mwRandomUnitPoint
 * (milkyway/src/milkyway_util.c:399-409). */
static void nbm_pick_shell(nbm_dsfmt* st, double rad, double* v)
{
    double rsq;
    do {
        v[0] = nbm_xrandom(st, -1.0, 1.0);
        v[1] = nbm_xrandom(st, -1.0, 1.0);
        v[2] = nbm_xrandom(st, -1.0, 1.0);
        rsq = v[2] * v[2] + (v[1] * v[1] + v[0] * v[0]);
    } while (rsq > 1.0);
    double rsc = rad / sqrt(rsq);
    v[0] *= rsc;
    v[1] *= rsc;
    v[2] *= rsc;
}

/* plummerRandomR (:47-56) */
static double nbm_plummer_r(nbm_dsfmt* st)
{
    double rnd = nbm_dsfmt_close_open(st);
    return 1.0 / sqrt(pow(rnd, -2.0 / 3.0) - 1.0);
}

/* plummerSelectFromG (:58-70) */
static double nbm_plummer_g(nbm_dsfmt* st)
{
    double x, y;
    do {
        x = nbm_xrandom(st, 0.0, 1.0);
        y = nbm_xrandom(st, 0.0, 0.1);
    } while (y > (x * x) * pow(1.0 - (x * x), 3.5));
    return x;
}

int nbm_plummer(nbm_dsfmt* prng, int nbody, double mass, double radius_scale,
                const double* rshift, const double* vshift, int ignore,
                double* pos, double* vel, double* body_mass, int* type)
{
    if (nbody <= 0) return nbody == 0 ? 0 : -1;
    const double zero[3] = { 0.0, 0.0, 0.0 };
    if (!rshift) rshift = zero;
    if (!vshift) vshift = zero;
    const double vel_scale = sqrt(mass / radius_scale);
    const double m = mass / nbody;
    for (int i = 0; i < nbody; ++i) {
        double r;
        do {
            r = nbm_plummer_r(prng);
        } while (isinf(r));
        double p[3], v[3];
        /* plummerBodyPosition (:72-80) */
        nbm_pick_shell(prng, radius_scale * r, p);
        /* plummerBodyVelocity (:82-92) with plummerRandomV (:72-79) */
        double x = nbm_plummer_g(prng);
        double vm = 1.4142135623730951 * x / sqrt(sqrt(1.0 + (r * r)));
        nbm_pick_shell(prng, vel_scale * vm, v);
        for (int d = 0; d < 3; ++d) {
            pos[3 * i + d] = p[d] + rshift[d];
            vel[3 * i + d] = v[d] + vshift[d];
        }
        body_mass[i] = m;
        if (type) type[i] = ignore ? -1 : 1;
    }
    return 0;
}
