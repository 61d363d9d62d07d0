/*
 * Pieces shared by the force walks (walk.hip, walk_group.hip): the launch
 * outputs, the closing half of a force evaluation and the FP64 compensated
 * sum that receives the FP32 partial sums.
 */
#ifndef NBHIP_WALK_H
#define NBHIP_WALK_H

#include "nbhip_device.h"

/* Output side of one walk launch. */
struct WalkOut {
    int n;
    int kick;
    int use_pot;
    int multi;          /* several simulations: lanes park at their own end */
    int all_eps2_pos;   /* every simulation has eps2 > 0 */
    int batch_shift;    /* mixed walk: log2 of the node block per FP32 partial sum */
    double far2;        /* group walk: squared far-field ratio (0: every cell per body) */
    double dtHalf;
    nbhip_potential pot;
};

/* Closing half of the force evaluation shared by both walks: external
 * potential (nbMapForceBody, nbody/src/nbody_grav.c:164-182), scatter to
 * the original body order, optional closing half-kick (advanceVelocities,
 * nbody/src/nbody_plain.c:301-315). */
__device__ __forceinline__ void nbh_walk_finish(const WalkOut& o, double px, double py,
                                                double pz, double ax, double ay, double az,
                                                unsigned j, double* __restrict__ oax,
                                                double* __restrict__ oay,
                                                double* __restrict__ oaz,
                                                double* __restrict__ vx, double* __restrict__ vy,
                                                double* __restrict__ vz)
{
    if (o.use_pot) {
        double ex, ey, ez;
        nbh_ext_accel(o.pot, px, py, pz, ex, ey, ez);
        ax = ax + ex;
        ay = ay + ey;
        az = az + ez;
    }
    oax[j] = ax;
    oay[j] = ay;
    oaz[j] = az;
    if (o.kick) {
        vx[j] = vx[j] + o.dtHalf * (ax + 0.0);
        vy[j] = vy[j] + o.dtHalf * (ay + 0.0);
        vz[j] = vz[j] + o.dtHalf * (az + 0.0);
    }
}

__device__ __forceinline__ void nbh_walk_counters(DevScalars* scal, unsigned nbb, unsigned nbc,
                                                  unsigned nop, unsigned steps)
{
    unsigned long long c0 = nbh_wave_sum_u64(nbb);
    unsigned long long c1 = nbh_wave_sum_u64(nbc);
    unsigned long long c2 = nbh_wave_sum_u64(nop);
    if (__lane_id() == 0) {
        atomicAdd(&scal->counters[0], c0);
        atomicAdd(&scal->counters[1], c1);
        atomicAdd(&scal->counters[2], c2);
        atomicAdd(&scal->counters[3], (unsigned long long) steps);
    }
}

/* Kahan-compensated FP64 accumulation of FP32 partial sums. */
__device__ __forceinline__ void nbh_kahan_add(double& sum, double& comp, double v)
{
    const double y = v - comp;
    const double t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

#endif
