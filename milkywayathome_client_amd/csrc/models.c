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

/* ---------------------------------------------------------------------- */
/* host external potential (same order as the device nbh_ext_accel)       */
/* ---------------------------------------------------------------------- */
static double nbm_pow1p5(double x)
{
    double s = sqrt(x);
    double e = fma(-s, s, x);
    double ds = e / (2.0 * s);
    double p = x * s;
    double pe = fma(x, s, -p);
    return p + (pe + x * ds);
}

static double nbm_cube(double t)
{
    double h = t * t;
    double l = fma(t, t, -h);
    double p = h * t;
    double pe = fma(h, t, -p);
    return p + (pe + l * t);
}

static void nbm_mn(double M, double da, double db, double x, double y, double z, double* t)
{
    double zp = sqrt(z * z + db * db);
    double azp = da + zp;
    double rp = x * x + y * y + azp * azp;
    double rth = nbm_pow1p5(rp);
    t[0] = -M * x / rth;
    t[1] = -M * y / rth;
    t[2] = -M * z * azp / (zp * rth);
}

void nbm_ext_accel(const double* P, const double* pos, double* acc)
{
    const double x = pos[0], y = pos[1], z = pos[2];
    const double limit = 0.00390625;
    double ab = sqrt(z * z + (y * y + x * x));
    double r = (double) (ab <= limit) * limit + (double) (ab > limit) * ab;
    double a[3] = { 0.0, 0.0, 0.0 }, t[3] = { 0.0, 0.0, 0.0 };
    if ((int) P[3] == 1) nbm_mn(P[4], P[5], P[6], x, y, z, a);
    t[0] = t[1] = t[2] = 0.0;
    if ((int) P[7] == 1) nbm_mn(P[8], P[9], P[10], x, y, z, t);
    a[0] += t[0]; a[1] += t[1]; a[2] += t[2];
    t[0] = t[1] = t[2] = 0.0;
    if ((int) P[11] == 1) {
        double v0 = P[12], d = P[13], q = P[14];
        double zq = z / q;
        double denom = d * d + x * x + y * y + zq * zq;
        double k = -2.0 * v0 * v0 / denom;
        t[0] = k * x; t[1] = k * y; t[2] = k * z / (q * q);
    } else if ((int) P[11] == 2) {
        double ha = P[13];
        double M = P[12] * P[12] * ha / 0.2162165954;
        double ar = ha + r;
        double c = (-M / (r * r)) * (log(ar / ha) / r - 1 / ar);
        t[0] = c * x; t[1] = c * y; t[2] = c * z;
    }
    a[0] += t[0]; a[1] += t[1]; a[2] += t[2];
    t[0] = t[1] = t[2] = 0.0;
    if ((int) P[0] == 1) {
        double tmp = P[2] + r;
        double s = -P[1] / (r * (tmp * tmp));
        t[0] = s * x; t[1] = s * y; t[2] = s * z;
    } else if ((int) P[0] == 2) {
        double tmp = sqrt(P[2] * P[2] + r * r);
        double s = -P[1] / nbm_cube(tmp);
        t[0] = s * x; t[1] = s * y; t[2] = s * z;
    }
    acc[0] = a[0] + t[0];
    acc[1] = a[1] + t[1];
    acc[2] = a[2] + t[2];
}

void nbm_lbr_to_cartesian(const double* lbr, double sun, double* xyz)
{
    const double l = lbr[0] * 3.14159265358979323846 / 180.0;
    const double b = lbr[1] * 3.14159265358979323846 / 180.0;
    const double r = lbr[2];
    xyz[0] = r * cos(l) * cos(b) - sun;
    xyz[1] = r * sin(l) * cos(b);
    xyz[2] = r * sin(b);
}

void nbm_reverse_orbit(const double* pot, const double* pos, const double* vel, double tstop,
                       double dt, double* fpos, double* fvel)
{
    double x[3] = { pos[0], pos[1], pos[2] };
    double v[3] = { -vel[0], -vel[1], -vel[2] };
    double a[3];
    const double dt_half = dt / 2.0;
    nbm_ext_accel(pot, x, a);
    for (double t = 0; t >= tstop * (-1); t -= dt) {
        for (int d = 0; d < 3; ++d) v[d] += dt_half * a[d];
        for (int d = 0; d < 3; ++d) x[d] += dt * v[d];
        nbm_ext_accel(pot, x, a);
        for (int d = 0; d < 3; ++d) v[d] += dt_half * a[d];
    }
    for (int d = 0; d < 3; ++d) {
        fpos[d] = x[d];
        fvel[d] = -v[d];
    }
}

/* two-component Plummer density and enclosed-mass integrand used for the
 * Henon length unit (nbody/src/nbody_virial.c:22-36) */
static double nbm_dwarf_density(double r, double ab, double ad, double mb, double md)
{
    double b = mb / pow(ab, 3.0) * pow(1.0 + r * r / ab / ab, -2.5);
    double d = md / pow(ad, 3.0) * pow(1.0 + r * r / ad / ad, -2.5);
    return 3.0 / 4.0 / 3.14159265358979323846 * (b + d);
}

static double nbm_dwarf_integral(double r, double ab, double ad, double mb, double md)
{
    double b = mb / ab * r * r * pow(1.0 + r * r / ab / ab, -0.5);
    double d = md / ad * r * r * pow(1.0 + r * r / ad / ad, -0.5);
    return 1.0 / 4.0 / 3.14159265358979323846 * (b + d);
}

double nbm_virial_radius(double a_b, double a_d, double m_b, double m_d)
{
    static const double w[5] = { 0.2369268850561891, 0.4786286704993665, 0.5688888888888889,
                                 0.4786286704993665, 0.2369268850561891 };
    static const double x[5] = { -0.9061798459386640, -0.5384693101056831, 0.0,
                                 0.5384693101056831, 0.9061798459386640 };
    const int n = 100000;
    const double upper = 1000.0 * (a_b > a_d ? a_b : a_d);
    const double width = (upper - 0.0) / (n * 1.0);
    double integral = 0.0;
    for (int i = 0; i < n; ++i) {
        double a1 = 0.0 + i * width;
        double b1 = 0.0 + (i + 1) * width;
        for (int j = 0; j < 5; ++j) {
            double r1 = width * x[j] / 2.0 + (a1 + b1) / 2.0;
            integral += w[j] * nbm_dwarf_density(r1, a_b, a_d, m_b, m_d) *
                        nbm_dwarf_integral(r1, a_b, a_d, m_b, m_d) * width / 2.0;
        }
    }
    const double mt = m_b + m_d;
    const double pi = 3.14159265358979323846;
    return mt * mt / 16.0 / pi / pi / integral;
}

double nbm_calculate_eps2(double nbody, double a_b, double a_d, double m_b, double m_d)
{
    double rv = nbm_virial_radius(a_b, a_d, m_b, m_d);
    double eps = rv * 0.98 * pow(nbody, -0.26);
    double eps2 = (eps * eps) / 1.0;
    if (eps2 <= 2.220446049250313e-16) eps2 = 2.220446049250313e-16;
    return eps2;
}

double nbm_calculate_timestep(double mass, double r0)
{
    const double pi_4_3 = 4.18879020478639098461685784437;
    return (1.0 / 10.0) * (1.0 / 10.0) * sqrt((pi_4_3 * (r0 * r0 * r0)) / mass);
}
