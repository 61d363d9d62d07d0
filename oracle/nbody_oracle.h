/*
 * CPU oracle for the Barnes-Hut force step and KDK leapfrog of the
 * Milkyway@Home nbody client.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference CPU algorithm, used by the test-suite, the smoke check and the
 * CPU-baseline leg of the benchmark as the *checker*.  The product path
 * (libnbhip) never links, loads or calls it.
 *
 * Every function cites the reference source it restates
 * (paths relative to the reference repository root).
 *
 * Pinning (what checks each function against something other than itself):
 *   orc_expand_box / orc_build_tree  - brute-force prefix enumeration on tiny
 *       N (cells == Morton prefixes shared by >= 2 massive bodies), exact
 *       dyadic midpoints, maxDepth/structure-error rules   (tests/test_oracle_tree.py)
 *   orc_tree cofm / rcrit2 / quad   - direct per-cell sums over the bodies of
 *       the cell (parallel-axis identity for Q), all three opening criteria
 *       computed from their closed forms                  (tests/test_oracle_tree.py)
 *   orc_grav_map (walk)             - theta = 0 equals the direct sum to
 *       rounding; far-field two-body and point-pair quadrupole closed forms;
 *       quadrupole error scaling                           (tests/test_oracle_walk.py)
 *   orc_ext_accel                   - minus the finite-difference gradient of
 *       each component's closed-form potential             (tests/test_oracle_potential.py)
 *   orc_direct_accel                - two-body closed form, momentum
 *       conservation (sum m a = 0)                         (tests/test_oracle_walk.py)
 *   orc_step / orc_run              - hand-computed KDK on one body in a
 *       uniform field, time reversibility, energy drift bound
 *                                                          (tests/test_oracle_integrator.py)
 * The reference itself cannot be compiled here from its own sources: its
 * nbody path includes cmake-generated headers (nbody_config.h,
 * milkyway_config.h) and links the missing nbody_bessel.c, so no
 * bit-for-bit reference binary pins the exact opening order.  That part of
 * the oracle is pinned by the theta = 0 / direct-sum identity and the
 * brute-force tree checks above.
 */
#ifndef NBODY_ORACLE_H
#define NBODY_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Opening criteria, numbered as criterion_t (nbody/include/nbody_types.h:166-173). */
enum { ORC_TREECODE = 0, ORC_SW93 = 1, ORC_BH86 = 2, ORC_EXACT = 3 };

/* Component codes, numbered as nbody/include/nbody_potential_types.h. */
enum { ORC_SPH_NONE = 0, ORC_SPH_HERNQUIST = 1, ORC_SPH_PLUMMER = 2 };
enum { ORC_DISK_NONE = 0, ORC_DISK_MN = 1 };
enum { ORC_HALO_NONE = 0, ORC_HALO_LOG = 1, ORC_HALO_NFW = 2 };

/* Status bits, valued as NBodyStatus (nbody/include/nbody_types.h:541-577). */
enum {
    ORC_SUCCESS = 0,
    ORC_TREE_STRUCTURE_ERROR = 1 << 3,
    ORC_TREE_INCEST_FATAL = 1 << 4,
    ORC_UNSUPPORTED = 1 << 11,
    ORC_TREE_INCEST_NONFATAL = 1 << 24
};

typedef struct {
    int sphere_type;  double sphere_mass, sphere_scale;
    int disk_type;    double disk_mass, disk_a, disk_b;
    int disk2_type;   double disk2_mass, disk2_a, disk2_b;
    int halo_type;    double halo_vhalo, halo_d, halo_q;
} orc_potential;

typedef struct {
    double eps2;
    double theta;
    double timestep;
    double tree_rsize;     /* initial root side (ctx->treeRSize) */
    int criterion;
    int use_quad;
    int allow_incest;
    int use_potential;     /* 0 = EXTERNAL_POTENTIAL_NONE, 1 = DEFAULT */
    orc_potential pot;
} orc_ctx;

/* Node reference: >= 0 body index, <= -2 cell (-(c + 2)), -1 NULL. */
typedef struct {
    double pos[3];      /* geometric midpoint while building, centre of mass after */
    double mass;
    double rcrit2;
    double quad[6];     /* xx xy xz yy yz zz */
    double mid[3];      /* geometric midpoint (kept for checks) */
    int sub[8];
    int more, next;
    int level;
} orc_cell;

typedef struct orc_tree orc_tree;

typedef struct {
    int n;
    double* pos;        /* 3n, x y z per body */
    double* vel;        /* 3n */
    double* mass;       /* n */
    double* acc;        /* 3n */
    double rsize;       /* persists across steps (grows only) */
    unsigned step;
    int tree_incest;
    orc_tree* tree;     /* tree of the last force evaluation */
} orc_state;

/* ---- tree ------------------------------------------------------------- */
orc_tree* orc_tree_new(void);
void orc_tree_free(orc_tree* t);
int orc_build_tree(const orc_ctx* ctx, int n, const double* pos, const double* mass,
                   double* rsize_io, orc_tree* t);
int orc_tree_ncell(const orc_tree* t);
int orc_tree_maxdepth(const orc_tree* t);
/* Pre-order (More/Next threading order) listing of every node.
 * kind[i] = 1 body / 0 cell, ref[i] = body index or cell index,
 * skip[i] = pre-order index of Next(node) (nnode if NULL),
 * level[i] = cell level (root 0) or -1 for bodies,
 * pos/mass/rcrit2/quad/mid are cell values (bodies: pos, mass, zeros). */
int orc_tree_nnode(const orc_tree* t);
void orc_tree_preorder(const orc_tree* t, const double* bpos, const double* bmass,
                       int* kind, int* ref, int* skip, int* level,
                       double* pos, double* mass, double* rcrit2, double* quad,
                       double* mid);

/* ---- forces ----------------------------------------------------------- */
void orc_ext_accel(const orc_potential* pot, const double* pos, double* acc);
int orc_grav_map(const orc_ctx* ctx, orc_state* st);
void orc_direct_accel(int n, const double* pos, const double* mass, double eps2,
                      int nsample, const int* idx, double* acc);
/* Walk only (no tree rebuild) for the bodies in idx, against the tree of the
 * last orc_grav_map; returns incest flag in *incest. */
void orc_walk_bodies(const orc_ctx* ctx, const orc_state* st, int nsample,
                     const int* idx, double* acc, int* incest);
/* Interaction counters of the last walk over idx (body-body, body-cell,
 * opened cells), for flop accounting. */
void orc_walk_counts(const orc_ctx* ctx, const orc_state* st, int nsample,
                     const int* idx, long long* counts3);

/* ---- integrator ------------------------------------------------------- */
orc_state* orc_state_new(int n, const double* pos, const double* vel, const double* mass,
                         double tree_rsize);
void orc_state_free(orc_state* st);
int orc_step(const orc_ctx* ctx, orc_state* st);
int orc_run(const orc_ctx* ctx, orc_state* st, unsigned nsteps);
void orc_set_threads(int nthreads);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif
