/*
 * nbmodels -- host-side initial conditions for the nbhip engine: the dSFMT
 * generator the reference uses for every model, and the Plummer sampler
 * behind `predefinedModels.plummer` (nbody/src/nbody_plummer.c).
 *
 * Plain C, no GPU.  Bodies are returned as n x 3 row-major arrays.
 */
#ifndef NBMODELS_H
#define NBMODELS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* dSFMT-19937 state (Saito & Matsumoto, "A PRNG specialized in double
 * precision floating point numbers", 2009), as created by DSFMT.create(seed)
 * (milkyway/src/milkyway_lua_dsfmt.c:51 -> dsfmt_init_gen_rand). */
#define NBM_DSFMT_N 191
typedef struct nbm_dsfmt {
    uint64_t status[2 * (NBM_DSFMT_N + 1)];
    int idx;
} nbm_dsfmt;

void nbm_dsfmt_init(nbm_dsfmt* st, uint32_t seed);
double nbm_dsfmt_close1_open2(nbm_dsfmt* st);   /* [1, 2) */
double nbm_dsfmt_close_open(nbm_dsfmt* st);     /* [0, 1) */
double nbm_dsfmt_open_close(nbm_dsfmt* st);     /* (0, 1] */
double nbm_dsfmt_open_open(nbm_dsfmt* st);      /* (0, 1) */
/* kind: 0 [1,2), 1 [0,1), 2 (0,1], 3 (0,1) */
void nbm_dsfmt_fill(nbm_dsfmt* st, int kind, int n, double* out);

/* nbGeneratePlummerCore (nbody/src/nbody_plummer.c:109-158): `nbody`
 * equal-mass bodies of a Plummer sphere of total `mass` and scale radius
 * `radius_scale`, shifted by rshift / vshift.  Draws from `prng`.  type[i] is
 * BODY(ignore) (nbody/include/nbody_types.h:91-92). */
int nbm_plummer(nbm_dsfmt* prng, int nbody, double mass, double radius_scale,
                const double* rshift, const double* vshift, int ignore,
                double* pos, double* vel, double* body_mass, int* type);

/* External potential for host-side orbit integration: the same components
 * and operation order as the device nbh_ext_accel (nbExtAcceleration,
 * nbody/src/nbody_potential.c:627-762).  Layout of pot[15]:
 * sphere_type, sphere_mass, sphere_scale, disk_type, disk_mass, disk_a,
 * disk_b, disk2_type, disk2_mass, disk2_a, disk2_b, halo_type, halo_vhalo,
 * halo_d, halo_q (types as doubles). */
void nbm_ext_accel(const double* pot15, const double* pos, double* acc);

/* lbrToCartesian (nbody/src/nbody_coordinates.c:51-71), l and b in degrees. */
void nbm_lbr_to_cartesian(const double* lbr, double sun_gc_dist, double* xyz);

/* nbReverseOrbit (nbody/src/nbody_orbit_integrator.c:44-86): test-particle
 * KDK backwards for tstop with step dt; returns the reversed-back state. */
void nbm_reverse_orbit(const double* pot15, const double* pos, const double* vel,
                       double tstop, double dt, double* final_pos, double* final_vel);

/* nbCalculateVirial (nbody/src/nbody_virial.c:38-102) and
 * nbCalculateEps2 (nbody/src/nbody_lua_models.c:158-169). */
double nbm_virial_radius(double a_b, double a_d, double m_b, double m_d);
double nbm_calculate_eps2(double nbody, double a_b, double a_d, double m_b, double m_d);
/* nbCalculateTimestep (nbody/src/nbody_lua_models.c:139-142) */
double nbm_calculate_timestep(double mass, double r0);

#ifdef __cplusplus
}
#endif

#endif
