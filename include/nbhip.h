/*
 * nbhip -- MI355X (gfx950) Barnes-Hut force step and KDK leapfrog for the
 * Milkyway@Home nbody client, exposed as a plain C ABI.
 *
 * The library replaces the OpenCL hot path of the reference
 * (nbody/src/nbody_cl.c + nbody/kernels/nbody_kernels.cl) behind the same
 * call sites.  It follows the CPU semantics of the reference (tree rooted at
 * the origin with a persistent, doubling root size; per-body opening
 * decisions; KDK leapfrog), not the OpenCL variants.
 *
 * All entry points return an NBodyStatus bit set (values identical to
 * NBodyStatus, nbody/include/nbody_types.h:541-577); 0 is success.
 * Arrays are plain host pointers; positions, velocities and accelerations
 * are n x 3 row-major (x, y, z per body).  The library never retains a
 * caller pointer past the call.
 */
#ifndef NBHIP_H
#define NBHIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status bits: NBodyStatus (nbody/include/nbody_types.h:541-577) ---- */
#define NBHIP_SUCCESS                0u
#define NBHIP_ERROR                  (1u << 1)
#define NBHIP_TREE_STRUCTURE_ERROR   (1u << 3)
#define NBHIP_TREE_INCEST_FATAL      (1u << 4)
#define NBHIP_CHECKPOINT_ERROR       (1u << 6)
#define NBHIP_GPU_ERROR              (1u << 7)   /* NBODY_CL_ERROR */
#define NBHIP_CAPABILITY_ERROR       (1u << 8)
#define NBHIP_UNSUPPORTED            (1u << 11)
#define NBHIP_USER_ERROR             (1u << 12)
#define NBHIP_MAX_DEPTH_ERROR        (1u << 16)
#define NBHIP_CELL_OVERFLOW_ERROR    (1u << 17)
#define NBHIP_TREE_INCEST_NONFATAL   (1u << 24)
#define NBHIP_STATUS_ALL_WARNINGS    0xFF000000u
#define NBHIP_STATUS_IS_FATAL(x)     (((x) & ~NBHIP_STATUS_ALL_WARNINGS) != 0u)

/* ---- opening criteria: criterion_t (nbody/include/nbody_types.h:166-173) ---- */
#define NBHIP_TREECODE 0
#define NBHIP_SW93     1
#define NBHIP_BH86     2
#define NBHIP_EXACT    3

/* ---- potential component codes (nbody/include/nbody_potential_types.h) ---- */
#define NBHIP_SPHERICAL_NONE      0
#define NBHIP_SPHERICAL_HERNQUIST 1
#define NBHIP_SPHERICAL_PLUMMER   2
#define NBHIP_DISK_NONE           0
#define NBHIP_DISK_MIYAMOTO_NAGAI 1
#define NBHIP_HALO_NONE           0
#define NBHIP_HALO_LOGARITHMIC    1
#define NBHIP_HALO_NFW            2

/* ---- walk arithmetic ---- */
#define NBHIP_PRECISION_FP64  0   /* plain form: FP64 end to end */
#define NBHIP_PRECISION_MIXED 1   /* FP32 walk, FP64 cross-batch sums, FP64 state */
/* as MIXED, with wave-coherent opening: a wavefront of 64 consecutive bodies
 * opens a cell when any of them must, so each body's interaction list
 * refines its reference list (more accurate, not the reference's lists) */
#define NBHIP_PRECISION_MIXED_COHERENT 2
/* LDS-staged group walk: each wavefront takes 64 consecutive bodies in
 * Morton order (cut into sub-groups where the order jumps between distant
 * cells), tests the tree's nodes 64 at a time against each sub-group's
 * bounding box with the context's criterion (a node is accepted when the
 * box's nearest point passes it, so every body's interaction list refines
 * its reference list), stages the accepted nodes in LDS as FP32 positions
 * relative to the sub-group's centre and runs a dense FP32 interaction loop
 * over them; each LDS batch is one FP32 partial sum added into FP64 with
 * Kahan compensation.  Groups run costliest first by a prediction from the
 * last walk, and a costly group's sub-groups may be shared out among
 * several waves; neither changes any result.  More accurate than the
 * reference lists; results depend on which bodies share a wavefront (not
 * bit-identical across batching). */
#define NBHIP_PRECISION_MIXED_GROUP 3

/* The parts of Potential (nbody/include/nbody_potential_types.h) that the
 * supported components read. */
typedef struct nbhip_potential {
    int sphere_type;  double sphere_mass, sphere_scale;
    int disk_type;    double disk_mass, disk_scale_length, disk_scale_height;
    int disk2_type;   double disk2_mass, disk2_scale_length, disk2_scale_height;
    int halo_type;    double halo_vhalo, halo_scale_length, halo_flatten_z;
} nbhip_potential;

/* The NBodyCtx fields on the force/integrate path
 * (nbody/include/nbody_types.h:450-521). */
typedef struct nbhip_ctx {
    double eps2;            /* ctx->eps2 */
    double theta;           /* ctx->theta */
    double timestep;        /* ctx->timestep */
    double tree_rsize;      /* ctx->treeRSize: initial root side, power of two */
    int criterion;          /* ctx->criterion */
    int use_quad;           /* ctx->useQuad */
    int allow_incest;       /* ctx->allowIncest */
    int use_potential;      /* ctx->potentialType: 0 NONE, 1 DEFAULT */
    int precision;          /* NBHIP_PRECISION_* */
    int lmc;                /* ctx->LMC: the LMC's moving potential and reflex shift
                               (nbody/src/nbody_plain.c:268-315,427-433,
                               nbody/src/nbody_grav.c:164-172) are not supported:
                               nonzero -> NBHIP_UNSUPPORTED */
    nbhip_potential pot;    /* ctx->pot */
} nbhip_ctx;

typedef struct nbhip_state nbhip_state;

/* Tree statistics of the last force evaluation. */
typedef struct nbhip_tree_info {
    double rsize;           /* st->tree.rsize */
    int ncell;              /* st->tree.cellUsed */
    int max_depth;          /* st->tree.maxDepth */
    int nnode;              /* cells + massive bodies */
    int nmassive;
    unsigned step;          /* st->step */
    unsigned status;        /* sticky status bits */
} nbhip_tree_info;

/* Replaces nbInitCL + nbInitNBodyStateCL (nbody/src/nbody_types.c:258-341):
 * validates the context (custom/unsupported potentials -> NBHIP_UNSUPPORTED,
 * as nbody_types.c:297-301), selects the device and reserves buffers for
 * `nbody` bodies. */
unsigned nbhip_create(const nbhip_ctx* ctx, int device, int nbody, nbhip_state** out);

/* Replaces nbMarshalBodies(st, TRUE) (nbody/src/nbody_cl.c:2211-2269):
 * host bodies -> device.  mass may contain zeros (test particles, never
 * inserted in the tree: nbody/src/nbody_tree.c:438-442). */
unsigned nbhip_upload(nbhip_state* st, const double* pos, const double* vel,
                      const double* mass);

/* Replaces nbMarshalBodies(st, FALSE): device -> host.  Any pointer may be
 * NULL.  acc receives st->acctab (the last force evaluation). */
unsigned nbhip_download(nbhip_state* st, double* pos, double* vel, double* acc);

/* Replaces nbGravMap (nbody/src/nbody_grav.c:305-328): tree build + walk +
 * external potential into the acceleration table, no integration. */
unsigned nbhip_gravity(nbhip_state* st);

/* Replaces nbStepSystem / nbStepSystemPlain (nbody/src/nbody.c:182-200,
 * nbody/src/nbody_plain.c:329-364): one KDK step, step counter += 1. */
unsigned nbhip_step(nbhip_state* st);

/* Replaces nbRunSystem / nbRunSystemPlain (nbody/src/nbody.c:202-212,
 * nbody/src/nbody_plain.c:366-461) without the output hooks: one force
 * evaluation, then `nsteps` KDK steps, stopping at the first fatal status. */
unsigned nbhip_run(nbhip_state* st, unsigned nsteps);

/* Runs `nsteps` KDK steps without the leading force evaluation (for callers
 * that already called nbhip_gravity, e.g. after upload).  Stops at the first
 * step with a fatal status: with graph steps the remaining replays do
 * nothing and the step counter counts only the steps that ran. */
unsigned nbhip_steps(nbhip_state* st, unsigned nsteps);

/* The host hooks of nbRunSystemPlain's loop (nbody/src/nbody_plain.c:366-461)
 * for nbhip_run_system.  Either callback may be NULL. */
typedef struct nbhip_run_hooks {
    unsigned nstep;             /* ctx->nStep: run until the step counter reaches it */
    int use_best_like;          /* ctx->useBestLike */
    double best_like_start;     /* ctx->BestLikeStart: first step fraction evaluated */
    double time_evolve;         /* ctx->timeEvolve (the best likelihood's time) */
    int keep_best_bodies;       /* keep the best step's bodies (nbf->outFileName) */
    unsigned chunk;             /* steps run between two checkpoint calls outside the
                                   evaluated steps (0 or 1: every step) */
    /* get_likelihood (nbody/src/nbody_plain.c:75-245) up to the comparison:
     * called after every step with step / nstep >= best_like_start; fills
     * lk[10] with nbSystemLikelihood's array (lk[0] the likelihood, the
     * components the caller does not use set to 0, as nbody_plain.c stores
     * them) from the state's current histograms (nbhip_histogram,
     * nbhip_histogram_dispersion).  Returns NBHIP_SUCCESS, or a status that
     * stops the run; a likelihood it could not form (no data file) is
     * reported as NaN and treated as the worst case. */
    unsigned (*likelihood)(void* user, nbhip_state* st, unsigned step, double* lk);
    /* nbCheckpoint (nbody/src/nbody_plain.c:60-73): called after every step
     * (every `chunk` steps outside the evaluated ones); writes a checkpoint
     * when it is time (nbhip_download + nbhip_tree_info_get give its
     * contents).  A nonzero return stops the run with that status. */
    unsigned (*checkpoint)(void* user, nbhip_state* st, unsigned step);
    void* user;
} nbhip_run_hooks;

/* st->bestLikelihood* after a run (nbody/src/nbody_plain.c:165-232). */
typedef struct nbhip_best_like {
    double lk[10];              /* the best step's array; lk[0] = bestLikelihood
                                   (DEFAULT_WORST_CASE 9999999.9 before any step) */
    double time;                /* bestLikelihood_time = step / nstep * timeEvolve */
    unsigned step;              /* the best step (0: none) */
    unsigned count;             /* bestLikelihood_count: improvements */
    unsigned evaluated;         /* steps whose likelihood was evaluated */
} nbhip_best_like;

/* Replaces nbRunSystem / nbRunSystemPlain (nbody/src/nbody.c:202-212,
 * nbody/src/nbody_plain.c:366-461) with its hooks: one force evaluation,
 * then KDK steps until the step counter reaches hooks->nstep.  With
 * use_best_like, every step from the first with step / nstep >=
 * best_like_start runs alone and is followed by the likelihood hook; the
 * likelihood is clamped as get_likelihood clamps it (beyond +-9999999.9 or
 * NaN: 9999999.9; exactly 0: 1e-9) and kept in *best when its absolute
 * value is smaller than the best so far (with keep_best_bodies the bodies
 * of that step too, for nbhip_best_bodies).  Steps before that run in
 * chunks (replayed graphs with nbhip_set_graph).  The checkpoint hook
 * follows every step (chunk).  The run stops at the first fatal status
 * (the likelihood of that step is still evaluated, as in the reference);
 * the final checkpoint (nbWriteFinalCheckpoint) is the caller's. */
unsigned nbhip_run_system(nbhip_state* st, const nbhip_run_hooks* hooks, nbhip_best_like* best);

/* The bodies (n x 3 each, upload order) of the best-likelihood step of the
 * last nbhip_run_system with keep_best_bodies (st->bestLikelihoodBodyTab);
 * NBHIP_USER_ERROR if none was kept.  Either pointer may be NULL. */
unsigned nbhip_best_bodies(nbhip_state* st, double* pos, double* vel);

unsigned nbhip_tree_info_get(nbhip_state* st, nbhip_tree_info* info);

/* Replaces nbReadCheckpointHeader's state restore (nbody/src/nbody_checkpoint.c:
 * 124-130: st->step, st->tree.rsize, st->treeIncest) for a GPU state resumed
 * from a checkpoint: after nbhip_upload (or nbhip_batch_upload) of the
 * checkpointed bodies, sets the step counter, the persistent root size of
 * simulation `sim` (0 for a single state; a power of two, as every size the
 * doubling rule produces from a power-of-two treeRSize) and ORs `status`
 * (sticky NBODY_* warning bits such as the incest flag) into its status.  The
 * next nbhip_run evaluates the force at the restored positions with the
 * restored root, so the resumed run continues bit for bit. */
unsigned nbhip_restore(nbhip_state* st, int sim, unsigned step, double rsize, unsigned status);

/* Pre-order node table of the last tree (cells and massive bodies in
 * More/Next threading order, nbody/src/nbody_tree.c:127-150).  Arrays hold
 * info.nnode entries; kind 0 cell / 1 body; ref = body index (original order)
 * or -1; skip = pre-order index of Next(); level = cell level or -1;
 * pos/mass = centre of mass (cells) or body; quad 6 per node (xx xy xz yy yz zz). */
unsigned nbhip_tree_download(nbhip_state* st, int* kind, int* ref, int* skip, int* level,
                             double* pos, double* mass, double* rcrit2, double* quad);

/* FP64 direct sum (nbGravity_Exact, nbody/src/nbody_grav.c:205-227) of the
 * self-gravity on the `nsample` bodies listed in idx, against all bodies. */
unsigned nbhip_direct_accel(nbhip_state* st, int nsample, const int* idx, double* acc);

/* Self-gravity only (no external potential) from the tree walk of the
 * current tree for all bodies, in the precision of the context. */
unsigned nbhip_walk_accel(nbhip_state* st, double* acc);

/* Walk counters of one extra counting walk over the current tree:
 * [0] body-body interactions, [1] body-cell interactions, [2] opened cells
 * (flop accounting), [3] wavefront node steps (64 x [3] lane slots were
 * issued; ([0]+[1]+[2]) / (64 x [3]) is the walk's lane utilisation). */
unsigned nbhip_walk_counters(nbhip_state* st, unsigned long long* counts4);
/* The same four counters and [4]: the group walk's far-field (cell, group)
 * expansions ([1] then counts only the cells evaluated per body). */
unsigned nbhip_walk_counters5(nbhip_state* st, unsigned long long* counts5);

/* Per-phase device time of the last step in milliseconds, measured with HIP
 * events on the library's stream: [0] drift+bbox, [1] keys+sort, [2] build,
 * [3] summarize, [4] walk (+potential, kick).  Requires profiling enabled. */
unsigned nbhip_set_profiling(nbhip_state* st, int enable);
unsigned nbhip_phase_times(nbhip_state* st, double* ms5);

/* ---- batches of independent simulations ----------------------------------
 * Many dwarfs of uneven body count, as a likelihood search evaluates them,
 * share one state: every phase of a step is one launch over all of them and
 * each simulation's results are bit-identical to its own single run.  No
 * reference counterpart: the reference runs one simulation per process
 * (nbody/src/nbody.c:399-560); these calls serve the batched search.
 * nbhip_gravity / nbhip_step / nbhip_steps / nbhip_run act on every
 * simulation and return the OR of their statuses. */
unsigned nbhip_batch_create(const nbhip_ctx* ctx, int device, int nsim, const int* nbody,
                            const double* eps2, nbhip_state** out);
unsigned nbhip_batch_upload(nbhip_state* st, int sim, const double* pos, const double* vel,
                            const double* mass);
unsigned nbhip_batch_download(nbhip_state* st, int sim, double* pos, double* vel,
                              double* acc);
/* Sticky status of each simulation (nsim entries). */
unsigned nbhip_batch_status(nbhip_state* st, unsigned* status);
unsigned nbhip_batch_tree_info(nbhip_state* st, int sim, nbhip_tree_info* info);
/* nbhip_tree_download for simulation `sim` of a batch: its own pre-order
 * segment (nbhip_batch_tree_info's nnode entries), with skip counted from
 * the segment's first node (its root) and ref from the simulation's first
 * body, so it compares directly with the simulation's tree built alone.
 * nbhip_tree_download returns the whole table of every simulation. */
unsigned nbhip_batch_tree_download(nbhip_state* st, int sim, int* kind, int* ref, int* skip,
                                   int* level, double* pos, double* mass, double* rcrit2,
                                   double* quad);

/* enable != 0: nbhip_steps captures one step into a hipGraph on first use and
 * replays it (no host synchronisation inside a step; the status is read
 * once per call). */
unsigned nbhip_set_graph(nbhip_state* st, int enable);

/* No counterpart in the reference.  Coherent and group mixed walks only
 * (NBHIP_PRECISION_MIXED_COHERENT, NBHIP_PRECISION_MIXED_GROUP): scale every
 * cell's squared critical radius by `scale` (>= 1) in the walk's records,
 * so the walk opens more and each body's interaction list refines its
 * reference list further (more accurate, more interactions).  Takes effect
 * at the next tree build; 1 is the default.  NBHIP_USER_ERROR for another
 * precision or a scale outside [1, 1e6]. */
unsigned nbhip_set_open_scale(nbhip_state* st, double scale);

/* No counterpart in the reference.  Group walk only
 * (NBHIP_PRECISION_MIXED_GROUP): an accepted cell whose distance |D| from a
 * group's centre satisfies u_max <= ratio * |D| (u_max: half the diagonal
 * of the group's bounding box) is evaluated once per group as a Taylor
 * expansion of its softened field about the centre -- field, Jacobian
 * (monopole + quadrupole) and third derivatives (monopole) -- instead of
 * once per body; the remainder is O(ratio^3) of the cell's field.  The
 * cells accepted do not change.  0 (the default) evaluates every accepted
 * cell per body.  Takes effect at the next walk; NBHIP_USER_ERROR for
 * another precision or a ratio outside [0, 0.5]. */
unsigned nbhip_set_group_expansion(nbhip_state* st, double ratio);

/* Walk launch parameters: workgroup size (64, 128, 256 threads) and
 * `batch`: for the mixed walks the node block (a power of two, in pre-order
 * node indices) over which a body's FP32 partial sum runs before the FP64
 * Kahan add -- it depends only on a body's own interaction list, so results
 * do not change with the workgroup size or with batching; for the group
 * walk the depth of its LDS interaction lists (64, 128 or 256 entries; one
 * list is one FP32 partial sum). */
typedef struct nbhip_tune {
    int block;
    int batch;
    double ms;              /* measured walk time of this setting */
} nbhip_tune;

/* Times every setting on the current tree (call after nbhip_gravity), keeps
 * the fastest for later steps and reports it in *best; trials (may be NULL)
 * receives up to max_trials measured settings, *ntrials their count. */
unsigned nbhip_autotune(nbhip_state* st, nbhip_tune* best, nbhip_tune* trials, int max_trials,
                        int* ntrials);
unsigned nbhip_set_tune(nbhip_state* st, const nbhip_tune* tune);

/* ---- histogram of the final state ---------------------------------------
 * HistogramParams (nbody/include/nbody_types.h) fields the lambda-beta
 * binning uses; angles in degrees, sun_gc_dist in kpc (NBodyCtx.sunGCDist). */
typedef struct nbhip_histogram_params {
    double phi, theta, psi;
    double lambda_start, lambda_end;
    double beta_start, beta_end;
    int lambda_bins, beta_bins;
    double sun_gc_dist;
} nbhip_histogram_params;

/* No counterpart in the reference.  nbhip_histogram without any device
 * allocation, for callers that cannot allocate device memory (code under
 * HIP graph capture, a best-likelihood loop that bins every step, serving
 * loops that forbid hipMalloc).  The counters live in a fixed device array
 * of NBHIP_NOALLOC_MAX_COUNTERS unsigned entries (4 MiB), one copy per GPU
 * held until the process ends; its address is looked up once per device by
 * nbhip_create / nbhip_batch_create.  nsim * (lambda_bins * beta_bins + 1)
 * must not exceed the maximum (NBHIP_USER_ERROR otherwise).  All calls on
 * one GPU share that array: calls that overlap on one GPU (several host
 * threads, several states) race.  Same outputs as nbhip_histogram. */
#define NBHIP_NOALLOC_MAX_COUNTERS (1u << 20)
unsigned nbhip_histogram_noalloc(nbhip_state* st, const nbhip_histogram_params* hp,
                                 unsigned* counts, unsigned* totals);

/* Per-body flag (n entries, all simulations, upload order); nonzero = the
 * body is left out of histograms, as ignoreBody leaves out the dark matter
 * of a mixed dwarf (nbody/src/nbody_histogram.c:863-865).  NULL clears. */
unsigned nbhip_set_ignore(nbhip_state* st, const unsigned char* ignore);

/* Raw counts of nbCreateHistogram (nbody/src/nbody_histogram.c:671-940) on
 * the current positions, computed on the GPU: counts[sim * bins + b] for
 * b = lambdaIndex * beta_bins + betaIndex, totals[sim] = bodies inside the
 * histogram (its totalNum); the caller normalises (count / totalNum) as
 * nbNormaliseHistogram does.  counts holds nsim * lambda_bins * beta_bins
 * entries, totals nsim (either may be NULL). */
unsigned nbhip_histogram(nbhip_state* st, const nbhip_histogram_params* hp, unsigned* counts,
                         unsigned* totals);

/* The settings of the per-bin histograms: outlier rejection (NBodyCtx's
 * IterMax, BetaSigma, VelSigma, DistSigma, PMSigma, BetaCorrect, VelCorrect,
 * DistCorrect, PMCorrect; nbody/src/nbody_lua_types/nbody_lua_nbodyctx.c:
 * 127-137) and the proper-motion frame (sunVelx/y/z in the model's velocity
 * units, NGPdec, NGPra, lNCP in radians; nbody/include/nbody_defaults.h:40-45). */
typedef struct nbhip_disp_params {
    int iter_max;
    double beta_sigma, vel_sigma, dist_sigma, pm_sigma;
    double beta_correct, vel_correct, dist_correct, pm_correct;
    double sun_vel[3];
    double ngp_dec, ngp_ra, l_ncp;
} nbhip_disp_params;

/* Histograms 1-7 of nbCreateHistogram (nbody/src/nbody_histogram.c:858-1110)
 * for simulation `sim` on the current state, after nbCalcDisp and iter_max
 * rounds of nbRemoveOutliers (nbody/src/nbody_mass.c:195-316): 1 the beta
 * dispersion, 2 the line-of-sight velocity dispersion (calc_vLOS,
 * nbody/src/nbody_mass.c:171-179), 3 the average line-of-sight velocity,
 * 4 the average beta, 5 the average heliocentric distance (calc_distance,
 * :181-187), 6 and 7 the average proper motions in declination and right
 * ascension (nbVXVYVZtomuDec / RA, nbody/src/nbody_coordinates.c:226-287).
 * out holds 14 * lambda_bins * beta_bins doubles:
 * out[(2 (k - 1) + 0) * bins + b] the value of histogram k in bin b
 * (lambda-major), out[(2 (k - 1) + 1) * bins + b] its error (-1 for an
 * average over 10 or fewer bodies).  Ignored bodies are left out as in
 * nbhip_histogram.  Device scratch (about 48 bytes per body of the
 * simulation plus 128 bytes per bin per 1024 bodies) is kept by the state,
 * grown when a call needs more and released by nbhip_release;
 * NBHIP_USER_ERROR when the per-workgroup partial sums would exceed 2^31
 * doubles. */
unsigned nbhip_histogram_dispersion(nbhip_state* st, int sim, const nbhip_histogram_params* hp,
                                    const nbhip_disp_params* dp, double* out);

/* ---- adjoint leapfrog ------------------------------------------------------
 * No counterpart in the reference.  Gradients of a loss on the final state
 * with respect to the initial positions, velocities, the masses and the
 * timestep, through nbRunSystem's KDK loop (nbody/src/nbody_plain.c:366-461),
 * with the tree of each step (cells, membership, opening decisions) held
 * fixed.  FP64 states only (NBODY_UNSUPPORTED otherwise).
 *
 * nbhip_adjoint_forward: from the uploaded state, one force evaluation and
 * nsteps KDK steps (the same results as nbhip_run), keeping a checkpoint of
 * the state every `every` steps in device memory.
 * nbhip_adjoint_backward: given dL/dpos and dL/dvel of the final state
 * (n x 3 each), recomputes each segment from its checkpoint and runs the
 * adjoint steps in reverse; writes dL/dpos0, dL/dvel0 (n x 3), dL/dmass (n)
 * and dL/dtimestep per simulation (nsim); any output may be NULL.  The state
 * is left as the forward run ended. */
unsigned nbhip_adjoint_forward(nbhip_state* st, unsigned nsteps, unsigned every);
/* Vector-Jacobian product of one force evaluation (nbhip_gravity) at the
 * current positions, with the current tree and its opening decisions held
 * fixed: given dL/dacc (n x 3) returns dL/dpos (n x 3) and dL/dmass (n),
 * either may be NULL.  The building block of nbhip_adjoint_backward. */
unsigned nbhip_gravity_vjp(nbhip_state* st, const double* gacc, double* gpos, double* gmass);
/* The reverse orbit of the dwarf's centre as a differentiable function of
 * the reversal time: nbReverseOrbit's KDK scheme
 * (nbody/src/nbody_orbit_integrator.c:44-86) with a fixed step count, step
 * t_rev / nsteps, in the external potential `pot` (the same expressions as
 * the force path's).  Writes the reversed end point (out_pos, out_vel, each
 * 3 doubles) and, by forward-mode differentiation through every step, its
 * derivative with respect to t_rev (dpos, dvel; either may be NULL).  Host
 * code (one body), no device needed.  No counterpart in the reference, whose
 * orbit is not differentiated. */
unsigned nbhip_orbit_tangent(const nbhip_potential* pot, const double* pos, const double* vel,
                             double t_rev, int nsteps, double* out_pos, double* out_vel,
                             double* dpos, double* dvel);
unsigned nbhip_adjoint_backward(nbhip_state* st, const double* gpos, const double* gvel,
                                double* gpos0, double* gvel0, double* gmass, double* gdt);

/* ---- checkpoint files (nbody/src/nbody_checkpoint.c) ----------------------
 * The reference's checkpoint layout (nbody_checkpoint.c:76-104 header,
 * nbFreezeState :488-536, nbThawState :396-486) for x86-64 hosts with real =
 * double: the "mwnbody" header (versions, nbody, step, sizeof(real),
 * sizeof(void*), nOrbitTrace, nShiftLMC, treeIncest, tree rsize, the raw
 * NBodyCtx), a size_t and the raw NBodyState, bodytab and
 * bestLikelihoodBodyTab as 96-byte Body records, the orbit trace and LMC
 * shifts as 32-byte mwvectors, "end"; the file is sized as the reference
 * sizes it (:213-214).  The NBodyCtx and NBodyState images are the host
 * program's and are written and returned byte for byte. */
typedef struct nbhip_checkpoint_header {
    unsigned major_version, minor_version;   /* NBODY_VERSION_MAJOR / _MINOR */
    unsigned nbody, step;                    /* st->nbody, st->step */
    unsigned real_size, ptr_size;            /* on read: sizeof(real), sizeof(void*) of the writer */
    unsigned n_orbit_trace, n_shift_lmc;     /* vectors that follow the bodies */
    unsigned tree_incest;                    /* st->treeIncest */
    double rsize;                            /* st->tree.rsize */
} nbhip_checkpoint_header;

typedef struct nbhip_checkpoint_io {
    unsigned major_version, minor_version;   /* written; on read the expected versions */
    void* ctx;                               /* NBodyCtx image (ctx_size bytes, required) */
    size_t ctx_size;                         /* sizeof(NBodyCtx): 592 on x86-64 */
    void* state;                             /* NBodyState image (state_size bytes) */
    size_t state_size;                       /* sizeof(NBodyState); on read 0 accepts any */
    short* type;                             /* Body.type per body (write: NULL gives 1, or -1 for
                                                bodies flagged by nbhip_set_ignore; read: filled
                                                when not NULL) */
    unsigned* id;                            /* Body.id per body (write: NULL gives the index) */
    double* best_pos;                        /* bestLikelihoodBodyTab positions, n x 3 (write:
                                                NULL gives the kept best-likelihood bodies of
                                                the state, else the current ones) */
    double* best_vel;                        /* ... and velocities */
    double* orbit_trace;                     /* mwvector x y z w per entry */
    unsigned n_orbit_trace;                  /* entries (read: capacity in, count out) */
    double* shift_lmc;                       /* mwvector per entry (LMC reflex accelerations) */
    unsigned n_shift_lmc;                    /* entries (read: capacity in, count out) */
    double lmc_pos[4], lmc_vel[4];           /* st->LMCpos, st->LMCvel (with shifts only) */
} nbhip_checkpoint_io;

/* Host-only halves (no device): nbFreezeState's file from host arrays
 * (pos, vel n x 3, mass n; hdr gives nbody, step, rsize, tree_incest; the
 * file is written to path.tmp<pid> and renamed over path, as
 * nbWriteCheckpointWithTmpFile does), and nbThawState's checks
 * (nbVerifyCheckpointHeader: "mwnbody", file size, real and pointer size,
 * version; the end marker) and contents; pos, vel, mass may be NULL to read
 * the header alone.  A file that fails a check -> NBHIP_CHECKPOINT_ERROR. */
unsigned nbhip_checkpoint_pack(const char* path, const nbhip_checkpoint_header* hdr,
                               const nbhip_checkpoint_io* io, const double* pos,
                               const double* vel, const double* mass);
unsigned nbhip_checkpoint_unpack(const char* path, nbhip_checkpoint_header* hdr,
                                 nbhip_checkpoint_io* io, double* pos, double* vel,
                                 double* mass);

/* nbWriteCheckpoint (nbody_checkpoint.c:629-672) for a GPU state: marshals
 * the bodies out (nbMarshalBodies, as nbCheckpointCL does,
 * nbody/src/nbody_cl.c:1549-1570) and writes them with the state's step,
 * root size and incest flag.  One simulation per file. */
unsigned nbhip_checkpoint_write(nbhip_state* st, const char* path, const nbhip_checkpoint_io* io);

/* nbReadCheckpoint (nbody_checkpoint.c:584-610) into a GPU state created for
 * the file's body count: uploads the bodies, restores step, root size and
 * incest flag (nbhip_restore) and the best-likelihood bodies
 * (nbhip_best_bodies), fills io's ctx / state images and optional arrays.
 * The next nbhip_run evaluates the force at the restored positions, as
 * nbRunSystemPlain does on resume (nbody/src/nbody_plain.c:388). */
unsigned nbhip_checkpoint_read(nbhip_state* st, const char* path, nbhip_checkpoint_io* io);

/* ---- two-component dwarf initial conditions ---------------------------------
 * predefinedModels.mixeddwarf (nbGenerateMixedDwarfCore,
 * nbody/src/nbody_mixeddwarf.c:765-1026) on the GPU: the reference's
 * component densities and potentials (nbody/src/nbody_dwarf_potential.c),
 * set_vars / get_extra_nfw_mass (:696-762: an NFW mass is M200, reset to the
 * mass inside 5 r200), radii by r_mag's rejection, speeds by rejection
 * against each component's Eddington distribution function in the total
 * potential below 0.99 v_esc, isotropic directions (velocity first), the
 * per-component centre-of-mass and momentum correction to pos / vel
 * (cm_correction_by_comp), baryons first.  The distribution functions are
 * tabulated once (Eddington's integral with analytic derivatives) and every
 * body draws from its own Philox4x32-10 stream (key = seed, counter = body
 * index and draw number): results do not depend on the launch shape and
 * repeat bit for bit, but they are not the reference's dSFMT stream. */
#define NBHIP_DWARF_PLUMMER       0   /* dwarf_t Plummer */
#define NBHIP_DWARF_NFW           1   /* dwarf_t NFW */
#define NBHIP_DWARF_GEN_HERNQUIST 2   /* dwarf_t General_Hernquist */

typedef struct nbhip_dwarf_params {
    int light_type, dark_type;            /* NBHIP_DWARF_* (comp1, comp2) */
    double light_mass, light_scale;       /* mass (an NFW mass is M200), scale radius */
    double dark_mass, dark_scale;
    unsigned nbody, nbody_baryon;         /* bodies; the first nbody_baryon are light */
    unsigned long long seed;              /* stream key */
    double pos[3], vel[3];                /* rShift, vShift: the orbit point */
} nbhip_dwarf_params;

/* Bodies into host arrays (pos, vel n x 3, mass n); comp_mass (2, may be
 * NULL) receives the components' masses after the NFW reset.  Light bodies
 * have type 1 and dark bodies -1 (BODY(isdark): left out of histograms).
 * NBHIP_ERROR if a body's rejection loop did not finish (4M trials) or a
 * speed exceeded its rejection envelope. */
unsigned nbhip_mixed_dwarf(const nbhip_dwarf_params* p, int device, double* pos, double* vel,
                           double* mass, double* comp_mass);
/* The same bodies written straight into simulation `sim` of a state (its
 * body count must be p->nbody), with the dark bodies flagged as ignored
 * (nbhip_set_ignore); nothing crosses PCIe.  Call nbhip_gravity next. */
unsigned nbhip_mixed_dwarf_state(nbhip_state* st, int sim, const nbhip_dwarf_params* p,
                                 double* comp_mass);
/* The tabulated distribution function of component comp (0 light, 1 dark)
 * at `count` energies (psi - v^2/2); NaN outside the table. */
unsigned nbhip_dwarf_df(const nbhip_dwarf_params* p, int device, int comp, int count,
                        const double* energy, double* f);
/* One Philox4x32-10 block (host code): out = philox(ctr[4], key[2]). */
unsigned nbhip_philox4x32_10(const unsigned* ctr, const unsigned* key, unsigned* out);

/* Waits for all work queued by the library. */
unsigned nbhip_sync(nbhip_state* st);

/* Replaces the OpenCL teardown inside destroyNBodyState
 * (nbody/src/nbody_types.c:142-175). */
void nbhip_release(nbhip_state* st);

/* Text of the last error of this thread (never NULL). */
const char* nbhip_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* NBHIP_H */
