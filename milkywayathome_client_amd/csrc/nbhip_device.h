/*
 * Device helpers shared by the nbhip kernels: fixed-point Morton keys,
 * the external Milky Way potential and small wave64 utilities.
 *
 * FP64 expressions keep the reference's operation order (the library is
 * compiled with -ffp-contract=off) so that the plain form reproduces the CPU
 * path bit for bit wherever the operations themselves are correctly rounded.
 */
#ifndef NBHIP_DEVICE_H
#define NBHIP_DEVICE_H

#include <hip/hip_runtime.h>
#include "nbhip_internal.h"

/* ---------------------------------------------------------------------- */
/* Fixed-point coordinates and Morton prefixes                            */
/* ---------------------------------------------------------------------- */
/* u = floor((x + R/2) * 2^63 / R), clamped to [0, 2^63 - 1].  With R a
 * power of two, x * 2^63/R is an exact scaling, so bit (62 - l) of u equals
 * the reference's comparison "midpoint <= x" at level l + 1 (nbSubIndex,
 * nbody/src/nbody_tree.c:36-52): every midpoint is an integer in this
 * fixed-point frame and the comparison is exact.  The root-size rule allows
 * |x| = R/2 exactly; x = +R/2 is on the upper side of every midpoint, so it
 * clamps to all ones (the clamp comes before the integer add, which would
 * overflow at t = 2^62), and x = -R/2 gives 0.  A NaN clamps to all ones. */
__host__ __device__ __forceinline__ unsigned long long nbh_fixed(double x, double key_scale)
{
    const double t = floor(x * key_scale);
    if (!(t < 4611686018427387904.0)) return 0x7FFFFFFFFFFFFFFFull;   /* t >= 2^62 */
    if (t < -4611686018427387904.0) return 0ull;
    return (unsigned long long) ((long long) t + (1LL << 62));
}

/* Spread the low 21 bits of v so that bit i lands on bit 3i. */
__device__ __forceinline__ unsigned long long nbh_spread21(unsigned long long v)
{
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x001F00000000FFFFull;
    v = (v | (v << 16)) & 0x001F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

/* Morton word of 21 levels; x is the most significant bit of each level
 * (octant index = 4*bx + 2*by + bz, nbody/src/nbody_tree.c:41-49). */
__device__ __forceinline__ unsigned long long nbh_morton21(unsigned long long bx,
                                                           unsigned long long by,
                                                           unsigned long long bz)
{
    return (nbh_spread21(bx) << 2) | (nbh_spread21(by) << 1) | nbh_spread21(bz);
}

/* Common Morton prefix, in bits (0..189), of two bodies given their
 * 63-bit fixed-point coordinates. */
__device__ __forceinline__ int nbh_prefix_bits(unsigned long long ax, unsigned long long ay,
                                               unsigned long long az, unsigned long long bx,
                                               unsigned long long by, unsigned long long bz)
{
    int ex = __clzll((long long) (ax ^ bx)) - 1;   /* bit 63 is always clear */
    int ey = __clzll((long long) (ay ^ by)) - 1;
    int ez = __clzll((long long) (az ^ bz)) - 1;
    int m = min(ex, min(ey, ez));
    if (m >= 63) return 189;
    return 3 * m + (ex == m ? 0 : (ey == m ? 1 : 2));
}

/* Number of octree levels shared by a key prefix of `bits` bits when the
 * top `sb` key bits hold the simulation id; -1 for none (different
 * simulations, or outside the array). */
__device__ __forceinline__ int nbh_levels(int bits, int sb)
{
    return bits < sb ? -1 : min(bits - sb, 189) / 3;
}

/* ---------------------------------------------------------------------- */
/* Correctly rounded powers used by the potentials (mw_pow -> crlibm        */
/* pow_rn, milkyway/include/milkyway_math_double.h:59): x^1.5 and t^3 are  */
/* formed in double-double and rounded once.                               */
/* ---------------------------------------------------------------------- */
__host__ __device__ __forceinline__ double nbh_pow1p5(double x)
{
    double s = sqrt(x);
    double e = fma(-s, s, x);
    double ds = e / (2.0 * s);
    double p = x * s;
    double pe = fma(x, s, -p);
    return p + (pe + x * ds);
}

__host__ __device__ __forceinline__ double nbh_cube(double t)
{
    double h = t * t;
    double l = fma(t, t, -h);
    double p = h * t;
    double pe = fma(h, t, -p);
    return p + (pe + l * t);
}

/* nbExtAcceleration (nbody/src/nbody_potential.c:627-762) restricted to the
 * supported components: Miyamoto-Nagai disk (:173-194), logarithmic halo
 * (:488-504), NFW halo (:506-515), Hernquist (:157-162) and Plummer
 * (:164-169) spheres. */
__host__ __device__ __forceinline__ void nbh_ext_accel(const nbhip_potential& pot, double x, double y,
                                              double z, double& oax, double& oay, double& oaz)
{
    const double limit = 0.00390625;
    double ab = sqrt(z * z + (y * y + x * x));
    double r = (double) (ab <= limit) * limit + (double) (ab > limit) * ab;
    double a0, a1, a2, t0, t1, t2;

    if (pot.disk_type == NBHIP_DISK_MIYAMOTO_NAGAI) {
        double M = pot.disk_mass, da = pot.disk_scale_length, db = pot.disk_scale_height;
        double zp = sqrt(z * z + db * db);
        double azp = da + zp;
        double rp = x * x + y * y + azp * azp;
        double rth = nbh_pow1p5(rp);
        a0 = -M * x / rth;
        a1 = -M * y / rth;
        a2 = -M * z * azp / (zp * rth);
    } else {
        a0 = a1 = a2 = 0.0;
    }
    if (pot.disk2_type == NBHIP_DISK_MIYAMOTO_NAGAI) {
        double M = pot.disk2_mass, da = pot.disk2_scale_length, db = pot.disk2_scale_height;
        double zp = sqrt(z * z + db * db);
        double azp = da + zp;
        double rp = x * x + y * y + azp * azp;
        double rth = nbh_pow1p5(rp);
        t0 = -M * x / rth;
        t1 = -M * y / rth;
        t2 = -M * z * azp / (zp * rth);
    } else {
        t0 = t1 = t2 = 0.0;
    }
    a0 += t0; a1 += t1; a2 += t2;

    if (pot.halo_type == NBHIP_HALO_LOGARITHMIC) {
        double v0 = pot.halo_vhalo, q = pot.halo_flatten_z, d = pot.halo_scale_length;
        double zq = z / q;
        double denom = d * d + x * x + y * y + zq * zq;
        double k = -2.0 * v0 * v0 / denom;
        t0 = k * x;
        t1 = k * y;
        t2 = k * z / (q * q);
    } else if (pot.halo_type == NBHIP_HALO_NFW) {
        double ha = pot.halo_scale_length;
        double M = pot.halo_vhalo * pot.halo_vhalo * ha / 0.2162165954;
        double ar = ha + r;
        double c = (-M / (r * r)) * (log(ar / ha) / r - 1 / ar);
        t0 = c * x; t1 = c * y; t2 = c * z;
    } else {
        t0 = t1 = t2 = 0.0;
    }
    a0 += t0; a1 += t1; a2 += t2;

    if (pot.sphere_type == NBHIP_SPHERICAL_HERNQUIST) {
        double tmp = pot.sphere_scale + r;
        double s = -pot.sphere_mass / (r * (tmp * tmp));
        t0 = s * x; t1 = s * y; t2 = s * z;
    } else if (pot.sphere_type == NBHIP_SPHERICAL_PLUMMER) {
        double sc = pot.sphere_scale;
        double tmp = sqrt(sc * sc + r * r);
        double s = -pot.sphere_mass / nbh_cube(tmp);
        t0 = s * x; t1 = s * y; t2 = s * z;
    } else {
        t0 = t1 = t2 = 0.0;
    }
    oax = a0 + t0;
    oay = a1 + t1;
    oaz = a2 + t2;
}


/* J g with J = d(nbh_ext_accel)/dx at (x, y, z), for the adjoint.  J is
 * symmetric (the acceleration is minus the gradient of a potential), so
 * this is also J^T g.  Same components and the same r clamp as
 * nbh_ext_accel: inside the clamp radius r is a constant.
 *   Miyamoto-Nagai:  a = -M u / D^1.5, u = (x, y, z (a + zeta) / zeta),
 *                    D = x^2 + y^2 + (a + zeta)^2, zeta = sqrt(z^2 + b^2),
 *                    J = -M (diag(1, 1, 1 + a b^2 / zeta^3) / D^1.5 - 3 u u^T / D^2.5)
 *   logarithmic:     a = k w, w = (x, y, z / q^2), k = -2 v0^2 / D,
 *                    D = d^2 + x^2 + y^2 + (z/q)^2, J = k diag(1, 1, 1/q^2) - 2 (k/D) w w^T
 *   radial f(r) x:   J = f I + f'(r) x x^T / r  (NFW, Hernquist, Plummer) */
__host__ __device__ __forceinline__ void nbh_mn_jac(double M, double da, double db, double x, double y,
                                           double z, double gx, double gy, double gz,
                                           double& ox, double& oy, double& oz)
{
    const double zeta = sqrt(z * z + db * db);
    const double azp = da + zeta;
    const double D = x * x + y * y + azp * azp;
    const double D15 = D * sqrt(D);
    const double D25 = D15 * D;
    const double uz = z * azp / zeta;
    const double ug = x * gx + y * gy + uz * gz;
    const double czz = 1.0 + da * db * db / (zeta * zeta * zeta);
    ox += -M * (gx / D15 - 3.0 * x * ug / D25);
    oy += -M * (gy / D15 - 3.0 * y * ug / D25);
    oz += -M * (czz * gz / D15 - 3.0 * uz * ug / D25);
}

__host__ __device__ __forceinline__ void nbh_ext_jac(const nbhip_potential& pot, double x, double y,
                                            double z, double gx, double gy, double gz,
                                            double& ox, double& oy, double& oz)
{
    const double limit = 0.00390625;
    const double ab = sqrt(z * z + (y * y + x * x));
    const bool clamp = ab <= limit;
    const double r = clamp ? limit : ab;
    const double xg = x * gx + y * gy + z * gz;
    ox = oy = oz = 0.0;
    if (pot.disk_type == NBHIP_DISK_MIYAMOTO_NAGAI)
        nbh_mn_jac(pot.disk_mass, pot.disk_scale_length, pot.disk_scale_height, x, y, z, gx, gy,
                   gz, ox, oy, oz);
    if (pot.disk2_type == NBHIP_DISK_MIYAMOTO_NAGAI)
        nbh_mn_jac(pot.disk2_mass, pot.disk2_scale_length, pot.disk2_scale_height, x, y, z, gx,
                   gy, gz, ox, oy, oz);
    double f = 0.0, fp = 0.0;          /* radial terms: f(r) and f'(r) */
    if (pot.halo_type == NBHIP_HALO_LOGARITHMIC) {
        const double v0 = pot.halo_vhalo, q = pot.halo_flatten_z, d = pot.halo_scale_length;
        const double q2 = q * q;
        const double zq = z / q;
        const double D = d * d + x * x + y * y + zq * zq;
        const double k = -2.0 * v0 * v0 / D;
        const double wz = z / q2;
        const double wg = x * gx + y * gy + wz * gz;
        const double c = 2.0 * k / D * wg;
        ox += k * gx - c * x;
        oy += k * gy - c * y;
        oz += k * gz / q2 - c * wz;
    } else if (pot.halo_type == NBHIP_HALO_NFW) {
        const double h = pot.halo_scale_length;
        const double M = pot.halo_vhalo * pot.halo_vhalo * h / 0.2162165954;
        const double hr = h + r;
        const double L = log(hr / h);
        const double r2 = r * r, r3 = r2 * r;
        f += -M * (L / r3 - 1.0 / (r2 * hr));
        fp += -M * (1.0 / (r3 * hr) - 3.0 * L / (r3 * r) + (2.0 * h + 3.0 * r) / (r3 * hr * hr));
    }
    if (pot.sphere_type == NBHIP_SPHERICAL_HERNQUIST) {
        const double a = pot.sphere_scale, M = pot.sphere_mass;
        const double ar = a + r;
        f += -M / (r * ar * ar);
        fp += M * (a + 3.0 * r) / (r * r * ar * ar * ar);
    } else if (pot.sphere_type == NBHIP_SPHERICAL_PLUMMER) {
        const double b = pot.sphere_scale, M = pot.sphere_mass;
        const double s2 = b * b + r * r;
        const double s = sqrt(s2);
        f += -M / (s2 * s);
        fp += 3.0 * M * r / (s2 * s2 * s);
    }
    const double c = clamp ? 0.0 : fp / r * xg;
    ox += f * gx + c * x;
    oy += f * gy + c * y;
    oz += f * gz + c * z;
}

/* ---------------------------------------------------------------------- */
/* wave64 helpers                                                          */
/* ---------------------------------------------------------------------- */
__device__ __forceinline__ unsigned long long nbh_lanemask_lt()
{
    unsigned lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ double nbh_wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ unsigned long long nbh_wave_sum_u64(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ unsigned nbh_wave_sum_u32(unsigned v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int nbh_wave_min_i(int v)
{
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ int nbh_wave_max_i(int v)
{
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

/* Lane set-up shared by both walks: the lane's simulation, its segment of
 * the node table and its start position; the wave's first and last node. */
struct LaneStart {
    int sim, self, lp, end, q0, qend;
    double eps2;
};

__device__ __forceinline__ LaneStart nbh_lane_start(int s, bool valid, int M,
                                                    const int* __restrict__ simid_s,
                                                    const SimTab* __restrict__ sims)
{
    LaneStart r;
    r.sim = valid ? simid_s[s] : 0;
    const SimTab& t = sims[r.sim];
    const bool live = valid && t.root >= 0;
    r.self = (valid && s < M) ? s : -2;
    r.lp = live ? t.root : INT_MAX;
    r.end = live ? t.end : 0;
    r.eps2 = t.eps2;
    r.q0 = __builtin_amdgcn_readfirstlane(nbh_wave_min_i(r.lp));
    r.qend = __builtin_amdgcn_readfirstlane(nbh_wave_max_i(r.end));
    return r;
}

__device__ __forceinline__ double nbh_wave_sum_f64(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

#endif
