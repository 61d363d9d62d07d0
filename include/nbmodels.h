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

#ifdef __cplusplus
}
#endif

#endif
