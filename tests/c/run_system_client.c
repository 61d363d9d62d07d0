/*
 * The drop-in seen from a C client: a plain C program, compiled by gcc
 * against include/nbhip.h only, that does what nbMain does with the GPU
 * path (nbody/src/nbody.c:417-497): create the state, marshal the bodies
 * in, run nbRunSystemPlain's loop with its hooks (best likelihood from
 * BestLikeStart on, checkpoints; nbody/src/nbody_plain.c:366-461) and
 * marshal the bodies out.
 *
 *   run_system_client cpu   context checks that need no device
 *   run_system_client gpu   the run itself, checked against
 *       - the oracle (oracle/liboracle.so, the C restatement of the
 *         reference CPU path): final positions and velocities bit for bit;
 *       - a second GPU state stepped one step at a time with nbhip_step and
 *         the same likelihood: best likelihood, its step, time, improvement
 *         count, evaluated steps, and the kept bodies, bit for bit;
 *       - the hook call sequence expected from BestLikeStart and the chunk.
 *
 * The likelihood hook bins the bodies on the GPU (nbhip_histogram, raw
 * counts of a 50 x 1 lambda-beta histogram), normalises the counts and
 * forms the EMD term of nbMatchEMD for one lambda range: the 1-D earth
 * mover's distance between the simulated and the data histogram (closed
 * form: the summed absolute difference of the cumulative fractions times
 * the bin width), then -300 log(1 - EMD / 50)
 * (nbody/src/nbody_emd.c:1278-1428).  The data histogram comes from a twin
 * dwarf with 2 % faster rotation, run 14 steps.
 *
 * Exit status 0 when every check passes; each failed check prints a line.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nbhip.h"
#include "nbmodels.h"
#include "nbody_oracle.h"

#define NB 5000
#define NSTEP 20
#define DT 1e-3
#define NLAMBDA 50
#define START 0.5
#define CHUNK 4

static int failures = 0;

#define CHECK(cond, ...)                                    \
    do {                                                    \
        if (!(cond)) {                                      \
            printf("FAIL %s:%d: ", __FILE__, __LINE__);     \
            printf(__VA_ARGS__);                            \
            printf("\n");                                   \
            ++failures;                                     \
        }                                                   \
    } while (0)

static nbhip_ctx make_ctx(void)
{
    nbhip_ctx c;
    memset(&c, 0, sizeof(c));
    c.eps2 = 1e-3;
    c.theta = 0.5;
    c.timestep = DT;
    c.tree_rsize = 4.0;
    c.criterion = NBHIP_BH86;
    c.use_quad = 1;
    c.allow_incest = 0;
    c.use_potential = 0;
    c.precision = NBHIP_PRECISION_FP64;
    return c;
}

/* A rotating ring of test dwarf bodies 26-34 kpc from the centre, the
 * same dSFMT stream for both rings (scale multiplies the rotation). */
static void ring(double vscale, double* pos, double* vel, double* mass)
{
    nbm_dsfmt prng;
    nbm_dsfmt_init(&prng, 3u);
    for (int i = 0; i < NB; ++i) {
        const double t = 2.0 * M_PI * nbm_dsfmt_close_open(&prng);
        const double r = 26.0 + 8.0 * nbm_dsfmt_close_open(&prng);
        const double z = 4.0 * (nbm_dsfmt_close_open(&prng) - 0.5);
        pos[3 * i + 0] = r * cos(t);
        pos[3 * i + 1] = r * sin(t);
        pos[3 * i + 2] = z;
        vel[3 * i + 0] = vscale * -200.0 * sin(t) + 10.0 * (nbm_dsfmt_close_open(&prng) - 0.5);
        vel[3 * i + 1] = vscale * 200.0 * cos(t) + 10.0 * (nbm_dsfmt_close_open(&prng) - 0.5);
        vel[3 * i + 2] = 10.0 * (nbm_dsfmt_close_open(&prng) - 0.5);
        mass[i] = 1e-3;
    }
}

/* HistogramParams of the sample workunits (nbody/sample_workunits/EMD_v172.lua) */
static nbhip_histogram_params hist_params(void)
{
    nbhip_histogram_params hp = { 128.79, 54.39, 90.70, -150.0, 150.0, -15.0, 15.0,
                                  NLAMBDA, 1, 8.0 };
    return hp;
}

static unsigned fractions(nbhip_state* st, double* frac)
{
    nbhip_histogram_params hp = hist_params();
    unsigned counts[NLAMBDA], total = 0;
    const unsigned rc = nbhip_histogram(st, &hp, counts, &total);
    if (rc) return rc;
    for (int b = 0; b < NLAMBDA; ++b) frac[b] = total ? (double) counts[b] / total : 0.0;
    return NBHIP_SUCCESS;
}

typedef struct client {
    double data[NLAMBDA];          /* the data histogram (fractions) */
    unsigned like_steps[NSTEP + 1];
    int nlike;
    unsigned ckpt_steps[NSTEP + 1];
    int nckpt;
} client;

static double emd_likelihood(const double* sim, const double* data)
{
    const double width = 300.0 / NLAMBDA;
    double cs = 0.0, cd = 0.0, emd = 0.0;
    for (int b = 0; b < NLAMBDA - 1; ++b) {
        cs += sim[b];
        cd += data[b];
        emd += fabs(cs - cd) * width;
    }
    return -300.0 * log(1.0 - emd / 50.0);
}

static unsigned like_hook(void* user, nbhip_state* st, unsigned step, double* lk)
{
    client* c = (client*) user;
    if (c->nlike <= NSTEP) c->like_steps[c->nlike++] = step;
    double sim[NLAMBDA];
    const unsigned rc = fractions(st, sim);
    if (rc) return rc;
    lk[0] = emd_likelihood(sim, c->data);
    for (int k = 1; k < 10; ++k) lk[k] = 0.0;
    lk[1] = lk[0];                 /* the EMD component */
    return NBHIP_SUCCESS;
}

static unsigned ckpt_hook(void* user, nbhip_state* st, unsigned step)
{
    (void) st;
    client* c = (client*) user;
    if (c->nckpt <= NSTEP) c->ckpt_steps[c->nckpt++] = step;
    return NBHIP_SUCCESS;
}

static int cpu_checks(void)
{
    nbhip_state* h = NULL;
    nbhip_ctx c = make_ctx();
    c.lmc = 1;
    CHECK(nbhip_create(&c, 0, 100, &h) == NBHIP_UNSUPPORTED && h == NULL,
          "an LMC context must be rejected with NBODY_UNSUPPORTED");
    c = make_ctx();
    c.tree_rsize = 3.0;
    CHECK(nbhip_create(&c, 0, 100, &h) == NBHIP_UNSUPPORTED,
          "treeRSize 3 must be rejected with NBODY_UNSUPPORTED");
    c = make_ctx();
    c.use_potential = 1;
    c.pot.halo_type = 7;
    CHECK(nbhip_create(&c, 0, 100, &h) == NBHIP_UNSUPPORTED,
          "an unknown halo must be rejected with NBODY_UNSUPPORTED");
    c = make_ctx();
    c.precision = 9;
    CHECK(nbhip_create(&c, 0, 100, &h) == NBHIP_USER_ERROR, "precision 9 is a user error");
    CHECK(nbhip_run_system(NULL, NULL, NULL) == NBHIP_USER_ERROR, "null state");
    return failures;
}

static int gpu_run(void)
{
    double* pos = malloc(3 * NB * sizeof(double));
    double* vel = malloc(3 * NB * sizeof(double));
    double* mass = malloc(NB * sizeof(double));
    double* p2 = malloc(3 * NB * sizeof(double));
    double* v2 = malloc(3 * NB * sizeof(double));
    double* bp = malloc(3 * NB * sizeof(double));
    double* bv = malloc(3 * NB * sizeof(double));
    double* keep_p = malloc(3 * NB * sizeof(double));
    double* keep_v = malloc(3 * NB * sizeof(double));
    const nbhip_ctx ctx = make_ctx();
    client cl;
    memset(&cl, 0, sizeof(cl));

    /* the data histogram: the faster twin after 14 steps */
    nbhip_state* twin = NULL;
    ring(1.02, pos, vel, mass);
    unsigned rc = nbhip_create(&ctx, 0, NB, &twin);
    if (rc) {
        printf("FAIL nbhip_create: %s\n", nbhip_last_error());
        return 1;
    }
    rc |= nbhip_upload(twin, pos, vel, mass);
    rc |= nbhip_run(twin, 14);
    rc |= fractions(twin, cl.data);
    nbhip_release(twin);
    CHECK(rc == 0, "twin run: %u", rc);

    /* the run: graph-replayed chunks before BestLikeStart, single steps after */
    ring(1.0, pos, vel, mass);
    nbhip_state* h = NULL;
    rc = nbhip_create(&ctx, 0, NB, &h);
    rc |= nbhip_upload(h, pos, vel, mass);
    rc |= nbhip_set_graph(h, 1);
    nbhip_run_hooks hooks;
    memset(&hooks, 0, sizeof(hooks));
    hooks.nstep = NSTEP;
    hooks.use_best_like = 1;
    hooks.best_like_start = START;
    hooks.time_evolve = NSTEP * DT;
    hooks.keep_best_bodies = 1;
    hooks.chunk = CHUNK;
    hooks.likelihood = like_hook;
    hooks.checkpoint = ckpt_hook;
    hooks.user = &cl;
    nbhip_best_like best;
    rc |= nbhip_run_system(h, &hooks, &best);
    CHECK(rc == 0, "nbhip_run_system: %u (%s)", rc, nbhip_last_error());
    nbhip_tree_info info;
    rc = nbhip_tree_info_get(h, &info);
    CHECK(rc == 0 && info.step == NSTEP, "step counter %u", info.step);
    rc = nbhip_download(h, p2, v2, NULL);
    CHECK(rc == 0, "download");
    rc = nbhip_best_bodies(h, bp, bv);
    CHECK(rc == 0, "best bodies");

    /* hook sequence: likelihood at every step with step / nstep >= start;
     * checkpoints after each chunk before that, then after every step */
    int ne = 0, nc = 0;
    unsigned exp_like[NSTEP + 1], exp_ckpt[NSTEP + 1];
    unsigned s = 0;
    while (s < NSTEP) {
        const int eval = (double) (s + 1) / NSTEP >= START;
        unsigned k = 1;
        if (!eval) {
            k = CHUNK;
            while (k > 1 && (double) (s + k) / NSTEP >= START) --k;
        }
        s += k;
        if (eval) exp_like[ne++] = s;
        exp_ckpt[nc++] = s;
    }
    CHECK(cl.nlike == ne && memcmp(cl.like_steps, exp_like, ne * sizeof(unsigned)) == 0,
          "likelihood hook at %d steps, expected %d", cl.nlike, ne);
    CHECK(cl.nckpt == nc && memcmp(cl.ckpt_steps, exp_ckpt, nc * sizeof(unsigned)) == 0,
          "checkpoint hook at %d steps, expected %d", cl.nckpt, nc);
    CHECK(best.evaluated == (unsigned) ne, "evaluated %u, expected %d", best.evaluated, ne);

    /* the same loop, one eager nbhip_step at a time */
    nbhip_state* g = NULL;
    rc = nbhip_create(&ctx, 0, NB, &g);
    rc |= nbhip_upload(g, pos, vel, mass);
    rc |= nbhip_gravity(g);
    double best_lk = 9999999.9, best_time = 0.0;
    unsigned best_step = 0, count = 0;
    for (unsigned step = 1; step <= NSTEP; ++step) {
        rc |= nbhip_step(g);
        if (!((double) step / NSTEP >= START)) continue;
        double frac[NLAMBDA];
        rc |= fractions(g, frac);
        double lk = emd_likelihood(frac, cl.data);
        if (lk > 9999999.9 || lk < -9999999.9 || isnan(lk)) lk = 9999999.9;
        if (lk == 0.0) lk = 1e-9;
        if (fabs(lk) < fabs(best_lk)) {
            best_lk = lk;
            best_step = step;
            best_time = ((double) step / NSTEP) * (NSTEP * DT);
            ++count;
            rc |= nbhip_download(g, keep_p, keep_v, NULL);
        }
    }
    CHECK(rc == 0, "step-by-step run: %u", rc);
    printf("best likelihood %.17g at step %u (time %.17g, %u improvements); "
           "step-by-step %.17g at step %u\n", best.lk[0], best.step, best.time, best.count,
           best_lk, best_step);
    CHECK(best.lk[0] == best_lk && best.step == best_step && best.count == count &&
              best.time == best_time,
          "best likelihood differs from the step-by-step loop");
    CHECK(best.step > 0 && best.step < NSTEP, "the data twin should give an interior best step");
    CHECK(memcmp(bp, keep_p, 3 * NB * sizeof(double)) == 0 &&
              memcmp(bv, keep_v, 3 * NB * sizeof(double)) == 0,
          "kept bodies differ from the step-by-step state at the best step");
    double* pg = malloc(3 * NB * sizeof(double));
    rc = nbhip_download(g, pg, NULL, NULL);
    CHECK(rc == 0 && memcmp(pg, p2, 3 * NB * sizeof(double)) == 0,
          "graph-chunked run and eager steps end in different places");
    free(pg);
    nbhip_release(g);
    nbhip_release(h);

    /* the oracle: the reference CPU path restated, same steps */
    orc_ctx oc;
    memset(&oc, 0, sizeof(oc));
    oc.eps2 = ctx.eps2;
    oc.theta = ctx.theta;
    oc.timestep = ctx.timestep;
    oc.tree_rsize = ctx.tree_rsize;
    oc.criterion = ORC_BH86;
    oc.use_quad = 1;
    orc_state* o = orc_state_new(NB, pos, vel, mass, ctx.tree_rsize);
    const int orc_rc = orc_run(&oc, o, NSTEP);
    CHECK(orc_rc == 0, "oracle run: %d", orc_rc);
    CHECK(memcmp(o->pos, p2, 3 * NB * sizeof(double)) == 0 &&
              memcmp(o->vel, v2, 3 * NB * sizeof(double)) == 0,
          "final bodies differ from the oracle's");
    double dmax = 0.0;
    for (int i = 0; i < 3 * NB; ++i) dmax = fmax(dmax, fabs(o->pos[i] - p2[i]));
    printf("max |x_gpu - x_oracle| after %d steps: %.3e\n", NSTEP, dmax);
    orc_state_free(o);

    free(pos); free(vel); free(mass); free(p2); free(v2); free(bp); free(bv);
    free(keep_p); free(keep_v);
    return failures;
}

int main(int argc, char** argv)
{
    if (argc != 2 || (strcmp(argv[1], "cpu") && strcmp(argv[1], "gpu"))) {
        fprintf(stderr, "usage: %s cpu|gpu\n", argv[0]);
        return 2;
    }
    const int f = strcmp(argv[1], "cpu") == 0 ? cpu_checks() : gpu_run();
    printf("%s: %d failure(s)\n", argv[1], f);
    return f ? 1 : 0;
}
