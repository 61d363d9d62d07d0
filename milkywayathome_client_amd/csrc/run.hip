/*
 * nbRunSystemPlain's loop with its host hooks (nbhip_run_system): the
 * best-likelihood evaluation of every late step and the checkpoint cadence
 * (nbody/src/nbody_plain.c:60-245,366-461), around the GPU steps.
 */
#include "nbhip_internal.h"

#include <cmath>
#include <cstring>

#define NBH_WORST_CASE 9999999.9   /* DEFAULT_WORST_CASE, nbody/include/nbody_defaults.h:77 */
#define NBH_BEST_CASE 1e-9         /* DEFAULT_BEST_CASE, :81 */

/* get_likelihood's clamp (nbody/src/nbody_plain.c:155-163) */
static double nbh_clamp_likelihood(double l)
{
    if (l > NBH_WORST_CASE || l < -NBH_WORST_CASE || std::isnan(l)) return NBH_WORST_CASE;
    if (l == 0.0) return NBH_BEST_CASE;
    return l;
}

/* First step s > from with s / nstep >= start, evaluated as the reference
 * does (curStep / Nstep >= ctx->BestLikeStart in double); nstep + 1 if none. */
static unsigned nbh_first_evaluated(unsigned from, unsigned nstep, double start)
{
    if (nstep == 0 || std::isnan(start)) return nstep + 1;
    double guess = std::ceil(start * (double) nstep);
    unsigned s = guess <= (double) from ? from + 1
                                        : (guess > (double) nstep ? nstep : (unsigned) guess);
    while (s > from + 1 && (double) (s - 1) / (double) nstep >= start) --s;
    while (s <= nstep && !((double) s / (double) nstep >= start)) ++s;
    return s;
}

static unsigned nbh_keep_best(nbhip_state* st)
{
    const size_t n = (size_t) st->n;
    if (!st->best_buf) NBH_CHECK(hipMalloc((void**) &st->best_buf, 6 * (n ? n : 1) * sizeof(double)));
    const double* src[6] = { st->px, st->py, st->pz, st->vx, st->vy, st->vz };
    for (int k = 0; k < 6; ++k)
        NBH_CHECK(hipMemcpyAsync(st->best_buf + k * n, src[k], n * sizeof(double),
                                 hipMemcpyDeviceToDevice, st->stream));
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_run_system(nbhip_state* st, const nbhip_run_hooks* h,
                                     nbhip_best_like* best)
{
    if (!st || !h) return NBHIP_USER_ERROR;
    nbhip_best_like b;
    memset(&b, 0, sizeof(b));
    b.lk[0] = NBH_WORST_CASE;                       /* st->bestLikelihood = DEFAULT_WORST_CASE */
    if (best) *best = b;
    st->best_kept = 0;

    unsigned rc = nbhip_gravity(st);                /* the force before the first step */
    if (NBHIP_STATUS_IS_FATAL(rc)) return rc;
    const unsigned nstep = h->nstep;
    const unsigned chunk = h->chunk > 1 ? h->chunk : 1;
    const unsigned first_eval = h->use_best_like
                                    ? nbh_first_evaluated(st->step, nstep, h->best_like_start)
                                    : nstep + 1;
    while (st->step < nstep) {
        const unsigned before = st->step;
        const bool eval = h->use_best_like && before + 1 >= first_eval;
        unsigned k = 1;
        if (!eval) {
            k = nstep - before;
            if (k > chunk) k = chunk;
            if (h->use_best_like && before + k >= first_eval) k = first_eval - 1 - before;
            if (k < 1) k = 1;
        }
        rc |= nbhip_steps(st, k);
        if (st->step == before) {
            /* nothing ran: a fatal status from an earlier call */
            if (!NBHIP_STATUS_IS_FATAL(rc)) rc |= NBHIP_ERROR;
            break;
        }
        if (eval && h->likelihood) {
            double lk[10];
            for (int i = 0; i < 10; ++i) lk[i] = NAN;
            const unsigned lr = h->likelihood(h->user, st, st->step, lk);
            if (lr) {
                rc |= lr;
                break;
            }
            const double like = nbh_clamp_likelihood(lk[0]);
            ++b.evaluated;
            if (std::fabs(like) < std::fabs(b.lk[0])) {
                memcpy(b.lk, lk, sizeof(b.lk));
                b.lk[0] = like;
                b.time = ((double) st->step / (double) nstep) * h->time_evolve;
                b.step = st->step;
                ++b.count;
                if (h->keep_best_bodies) {
                    const unsigned kr = nbh_keep_best(st);
                    if (kr) {
                        rc |= kr;
                        break;
                    }
                    st->best_kept = 1;
                }
            }
        }
        if (NBHIP_STATUS_IS_FATAL(rc)) break;
        if (h->checkpoint) {
            const unsigned cr = h->checkpoint(h->user, st, st->step);
            if (cr) {
                rc |= cr;
                break;
            }
        }
    }
    if (best) *best = b;
    return rc;
}

extern "C" unsigned nbhip_best_bodies(nbhip_state* st, double* pos, double* vel)
{
    if (!st) return NBHIP_USER_ERROR;
    if (!st->best_kept || !st->best_buf) {
        nbh_set_error("nbhip_best_bodies: no best-likelihood bodies were kept");
        return NBHIP_USER_ERROR;
    }
    NBH_CHECK(hipSetDevice(st->device));
    const size_t n = (size_t) st->n;
    if (n == 0) return NBHIP_SUCCESS;
    double* tmp = (double*) malloc(6 * n * sizeof(double));
    if (!tmp) return NBHIP_ERROR;
    hipError_t e = hipMemcpyAsync(tmp, st->best_buf, 6 * n * sizeof(double),
                                  hipMemcpyDeviceToHost, st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    if (e == hipSuccess) {
        for (size_t i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) {
                if (pos) pos[3 * i + d] = tmp[d * n + i];
                if (vel) vel[3 * i + d] = tmp[(3 + d) * n + i];
            }
    }
    free(tmp);
    if (e != hipSuccess) {
        nbh_set_error("nbhip_best_bodies: %s", hipGetErrorString(e));
        return NBHIP_GPU_ERROR;
    }
    return NBHIP_SUCCESS;
}
