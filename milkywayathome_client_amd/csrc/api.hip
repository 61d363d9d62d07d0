/*
 * C ABI of the nbhip library (declared in include/nbhip.h).
 *
 * One nbhip_state owns the device copy of one simulation and a HIP stream.
 * The sequence per KDK step mirrors nbStepSystemPlain
 * (nbody/src/nbody_plain.c:329-364): half-kick + drift (fused with the
 * bounding-box scan), root-size update, tree build, walk + external
 * potential + closing half-kick.
 */
#include "nbhip_internal.h"
#include "nbhip_device.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

static thread_local char g_err[512] = "";

void nbh_set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char* nbhip_last_error(void) { return g_err; }

static bool nbh_pow2(double r)
{
    if (!(r > 0.0) || !std::isfinite(r)) return false;
    int e;
    double f = std::frexp(r, &e);
    return f == 0.5;
}

static unsigned nbh_validate(const nbhip_ctx* c)
{
    if (c->criterion < NBHIP_TREECODE || c->criterion > NBHIP_EXACT) {
        nbh_set_error("invalid criterion %d", c->criterion);
        return NBHIP_USER_ERROR;
    }
    if (c->criterion != NBHIP_EXACT && !(c->theta >= 0.0)) {
        nbh_set_error("theta must be >= 0 for criterion != Exact");
        return NBHIP_USER_ERROR;
    }
    if (!nbh_pow2(c->tree_rsize)) {
        nbh_set_error("treeRSize %.17g is not a power of two: the exact fixed-point keys "
                      "need dyadic midpoints",
                      c->tree_rsize);
        return NBHIP_UNSUPPORTED;
    }
    const nbhip_potential& p = c->pot;
    if (c->use_potential) {
        if (p.sphere_type < 0 || p.sphere_type > NBHIP_SPHERICAL_PLUMMER ||
            p.disk_type < 0 || p.disk_type > NBHIP_DISK_MIYAMOTO_NAGAI ||
            p.disk2_type < 0 || p.disk2_type > NBHIP_DISK_MIYAMOTO_NAGAI ||
            p.halo_type < 0 || p.halo_type > NBHIP_HALO_NFW) {
            nbh_set_error("unsupported potential component");
            return NBHIP_UNSUPPORTED;
        }
    }
    if (c->lmc) {
        nbh_set_error("the LMC's moving potential (NBodyCtx.LMC) is not supported on the GPU");
        return NBHIP_UNSUPPORTED;
    }
    if (c->precision != NBHIP_PRECISION_FP64 && c->precision != NBHIP_PRECISION_MIXED &&
        c->precision != NBHIP_PRECISION_MIXED_COHERENT &&
        c->precision != NBHIP_PRECISION_MIXED_GROUP) {
        nbh_set_error("invalid precision %d", c->precision);
        return NBHIP_USER_ERROR;
    }
    return NBHIP_SUCCESS;
}

template <typename T>
static hipError_t nbh_alloc(T** p, size_t count)
{
    return hipMalloc((void**) p, (count > 0 ? count : 1) * sizeof(T));
}

static void nbh_adj_free(nbhip_state* st)
{
    NbhAdjoint* A = st->adj;
    if (!A) return;
    void* ptrs[] = { A->ckpt, A->ckpt_tab, A->seg, A->seg_tab, A->fin, A->fin_tab, A->g,
                     A->gdt, A->gnode };
    for (void* p : ptrs)
        if (p) (void) hipFree(p);
    free(A);
    st->adj = nullptr;
}

static void nbh_free_all(nbhip_state* st)
{
    void* ptrs[] = { st->px,   st->py,     st->pz,    st->vx,      st->vy,     st->vz,
                     st->m,    st->ax,     st->ay,    st->az,      st->simid,  st->key,
                     st->key_alt, st->perm, st->perm_alt, st->hist, st->dtot,  st->spos,
                     st->simid_s, st->su,  st->dadj,  st->cincl,   st->scan_part,
                     st->node,  st->nquad,  st->nmid,
                     st->nlevel, st->nparent, st->nsim_node, st->level_list, st->nodef,
                     st->scal, st->sims, st->ignore, st->hist_buf, st->disp_buf, st->pos4,
                     st->exact_part, st->nodeg, st->nkid, st->kbase, st->best_buf,
                     st->nrank, st->kid, st->blist, st->span, st->span_sorted, st->gcost,
                     st->gorder, st->bcost, st->gtask, st->nnext, st->bbox };
    for (void* p : ptrs)
        if (p) (void) hipFree(p);
    if (st->graph_exec) (void) hipGraphExecDestroy(st->graph_exec);
    if (st->graph) (void) hipGraphDestroy(st->graph);
    nbh_adj_free(st);
    free(st->h_body_off);
    free(st->h_body_n);
    free(st->h_sim_status);
}

static unsigned nbh_create(const nbhip_ctx* ctx, int device, int nsim, const int* nbody,
                           const double* eps2, nbhip_state** out)
{
    if (!ctx || !out || nsim < 1 || !nbody) {
        nbh_set_error("nbhip_create: bad arguments");
        return NBHIP_USER_ERROR;
    }
    *out = nullptr;
    long long total = 0;
    for (int k = 0; k < nsim; ++k) {
        if (nbody[k] < 0) {
            nbh_set_error("negative body count for simulation %d", k);
            return NBHIP_USER_ERROR;
        }
        total += nbody[k];
    }
    if (total > (1ll << 24)) {
        nbh_set_error("%lld bodies: at most 2^24 per state", total);
        return NBHIP_USER_ERROR;
    }
    if (nsim > 4096) {
        nbh_set_error("%d simulations: at most 4096 per state", nsim);
        return NBHIP_USER_ERROR;
    }
    unsigned rc = nbh_validate(ctx);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        nbh_set_error("no HIP device available");
        return NBHIP_CAPABILITY_ERROR;
    }
    if (device < 0 || device >= ndev) {
        nbh_set_error("device %d out of range (%d devices)", device, ndev);
        return NBHIP_CAPABILITY_ERROR;
    }
    NBH_CHECK(hipSetDevice(device));
    rc = nbh_scratch_init(device);
    if (rc) return rc;

    nbhip_state* st = (nbhip_state*) calloc(1, sizeof(nbhip_state));
    st->ctx = *ctx;
    st->device = device;
    st->n = (int) total;
    st->nsim = nsim;
    int sb = 0;
    while ((1 << sb) < nsim) ++sb;
    st->top_levels = (63 - sb) / 3;
    st->sim_shift = nsim > 1 ? 3 * st->top_levels : 64;
    st->node_cap = 5 * st->n + 1024 * nsim;
    st->open_scale = 1.0;
    st->group_far = 0.0;
    st->sort_blocks = st->n > 0 ? (st->n + NBH_SORT_TILE - 1) / NBH_SORT_TILE : 1;
    st->tune.block = 256;
    st->tune.batch = 1024;
    if (ctx->precision == NBHIP_PRECISION_MIXED_GROUP) {
        st->tune.block = 128;
        st->tune.batch = 128;
    }
    st->h_body_off = (int*) calloc(nsim, sizeof(int));
    st->h_body_n = (int*) calloc(nsim, sizeof(int));
    st->h_sim_status = (unsigned*) calloc(nsim, sizeof(unsigned));
    st->all_eps2_pos = 1;
    std::vector<SimTab> tab(nsim);
    std::vector<int> simid(st->n > 0 ? st->n : 1);
    int off = 0;
    for (int k = 0; k < nsim; ++k) {
        SimTab& t = tab[k];
        memset(&t, 0, sizeof(t));
        t.rsize = ctx->tree_rsize;
        t.key_scale = 9223372036854775808.0 / ctx->tree_rsize;
        t.eps2 = eps2 ? eps2[k] : ctx->eps2;
        t.body_off = off;
        t.body_n = nbody[k];
        t.first = t.root = t.end = -1;
        if (!(t.eps2 > 0.0)) st->all_eps2_pos = 0;
        st->h_body_off[k] = off;
        st->h_body_n[k] = nbody[k];
        st->exact_seg = std::max(st->exact_seg, nbh_exact_segments(nbody[k]));
        for (int i = 0; i < nbody[k]; ++i) simid[off + i] = k;
        off += nbody[k];
    }
    const size_t n = (size_t) st->n;
    const size_t nc = (size_t) st->node_cap;
    hipError_t e = hipSuccess;
#define A(ptr, cnt) if (e == hipSuccess) e = nbh_alloc(&st->ptr, (cnt))
    A(px, n); A(py, n); A(pz, n); A(vx, n); A(vy, n); A(vz, n); A(m, n);
    A(ax, n); A(ay, n); A(az, n); A(simid, n);
    A(key, n); A(key_alt, n); A(perm, n); A(perm_alt, n);
    A(hist, 256 * (size_t) st->sort_blocks);
    A(dtot, 256);
    A(pos4, n); A(spos, n); A(simid_s, n); A(su, 3 * n); A(dadj, nbh_dadj_cap(n)); A(cincl, n);
    A(scan_part, std::max<size_t>(4096, (nc + NBH_SCAN_TILE - 1) / NBH_SCAN_TILE));
    A(node, nc); A(nquad, 6 * nc); A(nmid, 3 * nc); A(nlevel, nc); A(nparent, nc);
    A(nsim_node, nc); A(level_list, nc); A(nnext, nc);
    if (ctx->precision == NBHIP_PRECISION_MIXED || ctx->precision == NBHIP_PRECISION_MIXED_COHERENT)
        A(nodef, nc + 1);   /* + 1: walk prefetch */
    if (ctx->precision == NBHIP_PRECISION_MIXED_GROUP) {
        const size_t ng = ((size_t) st->n + 63) / 64 + 1;
        A(nodeg, nc + (size_t) nsim);
        A(bcost, n ? n : 1);
        A(bbox, n ? n : 1);
        A(gcost, ng);
        A(gtask, 8 * ng);
        A(gorder, (size_t) nbh_group_tasks(st->n) + 1);
        /* the first walk has no costs yet: all zero, so every group in one bucket */
        if (e == hipSuccess) e = hipMemset(st->bcost, 0, (n ? n : 1) * sizeof(unsigned));
        if (e == hipSuccess) e = hipMemset(st->bbox, 0, (n ? n : 1) * sizeof(float));
    }
    A(nkid, nc); A(kbase, nc); A(nrank, nc); A(kid, nc); A(blist, nc); A(span, nc);
    A(span_sorted, nc);
    if (ctx->criterion == NBHIP_EXACT) A(exact_part, 4 * (size_t) st->exact_seg * n);
    A(scal, 1);
    A(sims, (size_t) nsim);
#undef A
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking);
    for (int k = 0; k < 6 && e == hipSuccess; ++k) e = hipEventCreate(&st->ev[k]);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&st->ev_counts, hipEventDisableTiming);
    if (e == hipSuccess)
        e = hipHostMalloc((void**) &st->h_counts_pinned, 64 * sizeof(int), hipHostMallocDefault);
    if (e != hipSuccess) {
        nbh_set_error("allocation for %d bodies failed: %s", st->n, hipGetErrorString(e));
        nbh_free_all(st);
        free(st);
        return NBHIP_GPU_ERROR;
    }
    DevScalars s0;
    memset(&s0, 0, sizeof(s0));
    s0.sim_bits = sb;
    NBH_CHECK(hipMemcpy(st->scal, &s0, sizeof(s0), hipMemcpyHostToDevice));
    NBH_CHECK(hipMemcpy(st->sims, tab.data(), nsim * sizeof(SimTab), hipMemcpyHostToDevice));
    if (n) {
        NBH_CHECK(hipMemcpy(st->simid, simid.data(), n * sizeof(int), hipMemcpyHostToDevice));
        NBH_CHECK(hipMemset(st->ax, 0, n * sizeof(double)));
        NBH_CHECK(hipMemset(st->ay, 0, n * sizeof(double)));
        NBH_CHECK(hipMemset(st->az, 0, n * sizeof(double)));
    }
    *out = st;
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_create(const nbhip_ctx* ctx, int device, int nbody, nbhip_state** out)
{
    return nbh_create(ctx, device, 1, &nbody, nullptr, out);
}

extern "C" unsigned nbhip_batch_create(const nbhip_ctx* ctx, int device, int nsim,
                                       const int* nbody, const double* eps2, nbhip_state** out)
{
    return nbh_create(ctx, device, nsim, nbody, eps2, out);
}

extern "C" void nbhip_release(nbhip_state* st)
{
    if (!st) return;
    (void) hipSetDevice(st->device);
    if (st->stream) (void) hipStreamSynchronize(st->stream);
    for (int k = 0; k < 6; ++k)
        if (st->ev[k]) (void) hipEventDestroy(st->ev[k]);
    if (st->ev_counts) (void) hipEventDestroy(st->ev_counts);
    if (st->h_counts_pinned) (void) hipHostFree(st->h_counts_pinned);
    if (st->stream) (void) hipStreamDestroy(st->stream);
    nbh_free_all(st);
    free(st);
}

static unsigned nbh_upload_range(nbhip_state* st, int off, int cnt, const double* pos,
                                 const double* vel, const double* mass)
{
    NBH_CHECK(hipSetDevice(st->device));
    const size_t n = (size_t) cnt;
    if (n == 0) return NBHIP_SUCCESS;
    std::vector<double> tmp(n);
    double* dst[6] = { st->px, st->py, st->pz, st->vx, st->vy, st->vz };
    for (int k = 0; k < 6; ++k) {
        const double* src = k < 3 ? pos : vel;
        const int d = k % 3;
        for (size_t i = 0; i < n; ++i) tmp[i] = src[3 * i + d];
        NBH_CHECK(hipMemcpyAsync(dst[k] + off, tmp.data(), n * sizeof(double),
                                 hipMemcpyHostToDevice, st->stream));
        NBH_CHECK(hipStreamSynchronize(st->stream));
    }
    NBH_CHECK(hipMemcpyAsync(st->m + off, mass, n * sizeof(double), hipMemcpyHostToDevice,
                             st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    st->have_tree = 0;
    return NBHIP_SUCCESS;
}

static unsigned nbh_download_range(nbhip_state* st, int off, int cnt, double* pos, double* vel,
                                   double* acc)
{
    NBH_CHECK(hipSetDevice(st->device));
    const size_t n = (size_t) cnt;
    if (n == 0) return NBHIP_SUCCESS;
    std::vector<double> tmp(n);
    double* outs[3] = { pos, vel, acc };
    double* srcs[3][3] = { { st->px, st->py, st->pz },
                           { st->vx, st->vy, st->vz },
                           { st->ax, st->ay, st->az } };
    for (int o = 0; o < 3; ++o) {
        if (!outs[o]) continue;
        for (int d = 0; d < 3; ++d) {
            NBH_CHECK(hipMemcpyAsync(tmp.data(), srcs[o][d] + off, n * sizeof(double),
                                     hipMemcpyDeviceToHost, st->stream));
            NBH_CHECK(hipStreamSynchronize(st->stream));
            for (size_t i = 0; i < n; ++i) outs[o][3 * i + d] = tmp[i];
        }
    }
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_upload(nbhip_state* st, const double* pos, const double* vel,
                                 const double* mass)
{
    if (!st || (st->n > 0 && (!pos || !vel || !mass))) {
        nbh_set_error("nbhip_upload: bad arguments");
        return NBHIP_USER_ERROR;
    }
    return nbh_upload_range(st, 0, st->n, pos, vel, mass);
}

extern "C" unsigned nbhip_download(nbhip_state* st, double* pos, double* vel, double* acc)
{
    if (!st) {
        nbh_set_error("nbhip_download: null state");
        return NBHIP_USER_ERROR;
    }
    return nbh_download_range(st, 0, st->n, pos, vel, acc);
}

extern "C" unsigned nbhip_batch_upload(nbhip_state* st, int sim, const double* pos,
                                       const double* vel, const double* mass)
{
    if (!st || sim < 0 || sim >= st->nsim ||
        (st->h_body_n[sim] > 0 && (!pos || !vel || !mass))) {
        nbh_set_error("nbhip_batch_upload: bad arguments");
        return NBHIP_USER_ERROR;
    }
    return nbh_upload_range(st, st->h_body_off[sim], st->h_body_n[sim], pos, vel, mass);
}

extern "C" unsigned nbhip_batch_download(nbhip_state* st, int sim, double* pos, double* vel,
                                         double* acc)
{
    if (!st || sim < 0 || sim >= st->nsim) {
        nbh_set_error("nbhip_batch_download: bad arguments");
        return NBHIP_USER_ERROR;
    }
    return nbh_download_range(st, st->h_body_off[sim], st->h_body_n[sim], pos, vel, acc);
}

/* Read back the device checks and fold them into the sticky statuses, mapping
 * a missed self-interaction onto the incest codes (nbIncestStatusCheck,
 * nbody/src/nbody_grav.c:295-303). */
static unsigned nbh_collect(nbhip_state* st)
{
    DevScalars s;
    std::vector<SimTab> tab(st->nsim);
    NBH_CHECK(hipMemcpyAsync(&s, st->scal, sizeof(s), hipMemcpyDeviceToHost, st->stream));
    NBH_CHECK(hipMemcpyAsync(tab.data(), st->sims, st->nsim * sizeof(SimTab),
                             hipMemcpyDeviceToHost, st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    /* st->status keeps what concerns the whole state (node capacity, launch
     * errors); h_sim_status[k] what one simulation's tree or walk reported. */
    st->status |= s.status;
    /* steps the device completed since the last read (a halted graph replay
     * does not count the steps it skipped) */
    st->step += s.steps_done - st->dev_steps;
    st->dev_steps = s.steps_done;
    unsigned rc = st->status;
    for (int k = 0; k < st->nsim; ++k) {
        unsigned f = tab[k].flags;
        if (tab[k].incest)
            f |= st->ctx.allow_incest ? NBHIP_TREE_INCEST_NONFATAL : NBHIP_TREE_INCEST_FATAL;
        st->h_sim_status[k] |= f;
        rc |= st->h_sim_status[k];
    }
    if (rc & NBHIP_CELL_OVERFLOW_ERROR)
        nbh_set_error("node table overflow: %d nodes for capacity %d", s.nnode, st->node_cap);
    else if (rc & NBHIP_TREE_STRUCTURE_ERROR)
        nbh_set_error("tree structure error (max depth %d)", s.maxdepth);
    else if (rc & NBHIP_ERROR)
        nbh_set_error("internal consistency error in the tree or walk");
    return rc;
}

static unsigned nbh_status_all(const nbhip_state* st)
{
    unsigned rc = st->status;
    for (int k = 0; k < st->nsim; ++k) rc |= st->h_sim_status[k];
    return rc;
}

static void nbh_mark(nbhip_state* st, int k)
{
    if (st->profiling) (void) hipEventRecord(st->ev[k], st->stream);
}

/* one force evaluation; drift = 1 prepends advancePosVel, kick = 1 appends
 * advanceVelocities */
static unsigned nbh_force(nbhip_state* st, int drift, int kick)
{
    nbh_mark(st, 0);
    if (st->ctx.criterion == NBHIP_EXACT) {
        nbh_bbox(st, drift);
        nbh_mark(st, 1);
        nbh_mark(st, 2);
        nbh_mark(st, 3);
        nbh_mark(st, 4);
    } else {
        nbh_bbox(st, drift);
        nbh_mark(st, 1);
        nbh_build(st, 0);
    }
    nbh_walk(st, kick, 0, st->ax, st->ay, st->az, 0, nullptr);
    nbh_mark(st, 5);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        nbh_set_error("kernel launch failed: %s", hipGetErrorString(e));
        return NBHIP_GPU_ERROR;
    }
    st->have_tree = st->ctx.criterion != NBHIP_EXACT;
    return NBHIP_SUCCESS;
}

static void nbh_phase_read(nbhip_state* st)
{
    if (!st->profiling) return;
    (void) hipEventSynchronize(st->ev[5]);
    for (int k = 0; k < 5; ++k) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, st->ev[k], st->ev[k + 1]) != hipSuccess) ms = 0.0f;
        st->phase_ms[k] = ms;
    }
}

extern "C" unsigned nbhip_gravity(nbhip_state* st)
{
    if (!st) return NBHIP_USER_ERROR;
    NBH_CHECK(hipSetDevice(st->device));
    unsigned rc = nbh_force(st, 0, 0);
    if (rc) return rc;
    rc = nbh_collect(st);
    nbh_phase_read(st);
    return rc;
}

extern "C" unsigned nbhip_step(nbhip_state* st)
{
    if (!st) return NBHIP_USER_ERROR;
    NBH_CHECK(hipSetDevice(st->device));
    unsigned rc = nbh_force(st, 1, 1);
    if (rc) return rc;
    nbh_step_end(st);
    rc = nbh_collect(st);
    nbh_phase_read(st);
    return rc;
}

/* One KDK step as a captured hipGraph: no host synchronisation inside. */
static unsigned nbh_capture_step(nbhip_state* st)
{
    if (st->graph_exec) (void) hipGraphExecDestroy(st->graph_exec);
    if (st->graph) (void) hipGraphDestroy(st->graph);
    st->graph_exec = nullptr;
    st->graph = nullptr;
    const int prof = st->profiling;
    st->profiling = 0;
    NBH_CHECK(hipStreamBeginCapture(st->stream, hipStreamCaptureModeThreadLocal));
    nbh_bbox(st, 1);
    if (st->ctx.criterion != NBHIP_EXACT) nbh_build(st, 1);
    nbh_walk(st, 1, 0, st->ax, st->ay, st->az, 0, nullptr);
    nbh_step_end(st);
    hipError_t e = hipStreamEndCapture(st->stream, &st->graph);
    st->profiling = prof;
    if (e != hipSuccess) {
        nbh_set_error("graph capture failed: %s", hipGetErrorString(e));
        return NBHIP_GPU_ERROR;
    }
    NBH_CHECK(hipGraphInstantiate(&st->graph_exec, st->graph, nullptr, nullptr, 0));
    st->graph_ready = 1;
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_steps(nbhip_state* st, unsigned nsteps)
{
    if (!st) return NBHIP_USER_ERROR;
    unsigned rc = nbh_status_all(st);
    if (NBHIP_STATUS_IS_FATAL(rc)) return rc;
    if (st->use_graph && nsteps > 0) {
        NBH_CHECK(hipSetDevice(st->device));
        if (!st->graph_ready) {
            rc = nbh_capture_step(st);
            if (rc) return rc;
        }
        for (unsigned k = 0; k < nsteps; ++k) NBH_CHECK(hipGraphLaunch(st->graph_exec, st->stream));
        st->have_tree = st->ctx.criterion != NBHIP_EXACT;
        return nbh_collect(st);
    }
    for (unsigned k = 0; k < nsteps; ++k) {
        rc |= nbhip_step(st);
        if (NBHIP_STATUS_IS_FATAL(rc)) return rc;
    }
    return rc;
}

extern "C" unsigned nbhip_set_graph(nbhip_state* st, int enable)
{
    if (!st) return NBHIP_USER_ERROR;
    st->use_graph = enable ? 1 : 0;
    st->graph_ready = 0;
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_run(nbhip_state* st, unsigned nsteps)
{
    if (!st) return NBHIP_USER_ERROR;
    unsigned rc = nbhip_gravity(st);
    if (NBHIP_STATUS_IS_FATAL(rc)) return rc;
    return rc | nbhip_steps(st, nsteps);
}

extern "C" unsigned nbhip_sync(nbhip_state* st)
{
    if (!st) return NBHIP_USER_ERROR;
    NBH_CHECK(hipSetDevice(st->device));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    return NBHIP_SUCCESS;
}

static unsigned nbh_info(nbhip_state* st, int sim, nbhip_tree_info* info)
{
    NBH_CHECK(hipSetDevice(st->device));
    DevScalars s;
    std::vector<SimTab> tab(st->nsim);
    NBH_CHECK(hipMemcpyAsync(&s, st->scal, sizeof(s), hipMemcpyDeviceToHost, st->stream));
    NBH_CHECK(hipMemcpyAsync(tab.data(), st->sims, st->nsim * sizeof(SimTab),
                             hipMemcpyDeviceToHost, st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    const SimTab& t = tab[sim];
    info->rsize = t.rsize;
    info->max_depth = t.maxdepth;
    info->step = st->step;
    info->status = st->h_sim_status[sim] | st->status;
    if (st->nsim == 1) {
        info->ncell = s.ncell;
        info->nnode = s.nnode;
        info->nmassive = s.nmassive;
        return NBHIP_SUCCESS;
    }
    int next_first = s.nmassive;
    for (int k = sim + 1; k < st->nsim; ++k)
        if (tab[k].first >= 0) {
            next_first = tab[k].first;
            break;
        }
    info->nmassive = t.first >= 0 ? next_first - t.first : 0;
    info->nnode = t.root >= 0 ? t.end - t.root : 0;
    info->ncell = info->nnode - info->nmassive;
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_tree_info_get(nbhip_state* st, nbhip_tree_info* info)
{
    if (!st || !info) return NBHIP_USER_ERROR;
    return nbh_info(st, 0, info);
}

extern "C" unsigned nbhip_restore(nbhip_state* st, int sim, unsigned step, double rsize,
                                  unsigned status)
{
    if (!st || sim < 0 || sim >= st->nsim) return NBHIP_USER_ERROR;
    int ex = 0;
    if (!(rsize > 0.0) || !std::isfinite(rsize) || std::frexp(rsize, &ex) != 0.5) {
        nbh_set_error("nbhip_restore: root size %.17g is not a positive power of two", rsize);
        return NBHIP_USER_ERROR;
    }
    NBH_CHECK(hipSetDevice(st->device));
    SimTab t;
    NBH_CHECK(hipMemcpyAsync(&t, st->sims + sim, sizeof(t), hipMemcpyDeviceToHost, st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    t.rsize = rsize;
    t.key_scale = 9223372036854775808.0 / rsize;
    NBH_CHECK(hipMemcpyAsync(st->sims + sim, &t, sizeof(t), hipMemcpyHostToDevice, st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    st->step = step;
    st->h_sim_status[sim] |= status;
    st->have_tree = 0;        /* a captured graph step reads the root from the table */
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_batch_tree_info(nbhip_state* st, int sim, nbhip_tree_info* info)
{
    if (!st || !info || sim < 0 || sim >= st->nsim) return NBHIP_USER_ERROR;
    return nbh_info(st, sim, info);
}

extern "C" unsigned nbhip_batch_status(nbhip_state* st, unsigned* status)
{
    if (!st || !status) return NBHIP_USER_ERROR;
    for (int k = 0; k < st->nsim; ++k) status[k] = st->h_sim_status[k] | st->status;
    return NBHIP_SUCCESS;
}

/* Node table segment [first, first + nn) of the last tree; skip and ref are
 * made relative to the segment and to body_off. */
static unsigned nbh_tree_segment(nbhip_state* st, int first, int nn, int body_off, int* kind,
                                 int* ref, int* skip, int* level, double* pos, double* mass,
                                 double* rcrit2, double* quad)
{
    if (nn <= 0) return NBHIP_SUCCESS;
    std::vector<NodeW> nodes(nn);
    std::vector<double> q(6 * (size_t) nn);
    std::vector<int> lev(nn);
    std::vector<unsigned> perm(st->n > 0 ? st->n : 1);
    NBH_CHECK(hipMemcpy(nodes.data(), st->node + first, nn * sizeof(NodeW),
                        hipMemcpyDeviceToHost));
    NBH_CHECK(hipMemcpy(q.data(), st->nquad + 6 * (size_t) first, 6 * (size_t) nn * sizeof(double),
                        hipMemcpyDeviceToHost));
    NBH_CHECK(hipMemcpy(lev.data(), st->nlevel + first, nn * sizeof(int), hipMemcpyDeviceToHost));
    if (st->n > 0)
        NBH_CHECK(hipMemcpy(perm.data(), st->perm, st->n * sizeof(unsigned),
                            hipMemcpyDeviceToHost));
    for (int i = 0; i < nn; ++i) {
        const NodeW& w = nodes[i];
        const bool body = w.body >= 0;
        if (kind) kind[i] = body ? 1 : 0;
        if (ref) ref[i] = body ? (int) perm[w.body] - body_off : -1;
        if (skip) skip[i] = w.next - first;
        if (level) level[i] = body ? -1 : lev[i];
        if (pos) { pos[3 * i] = w.x; pos[3 * i + 1] = w.y; pos[3 * i + 2] = w.z; }
        if (mass) mass[i] = w.m;
        if (rcrit2) rcrit2[i] = body ? 0.0 : w.rcrit2;
        if (quad)
            for (int k = 0; k < 6; ++k) quad[6 * i + k] = body ? 0.0 : q[6 * (size_t) i + k];
    }
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_tree_download(nbhip_state* st, int* kind, int* ref, int* skip,
                                        int* level, double* pos, double* mass, double* rcrit2,
                                        double* quad)
{
    if (!st) return NBHIP_USER_ERROR;
    if (!st->have_tree) {
        nbh_set_error("no tree: call nbhip_gravity first (and criterion != Exact)");
        return NBHIP_USER_ERROR;
    }
    NBH_CHECK(hipSetDevice(st->device));
    DevScalars s;
    NBH_CHECK(hipMemcpyAsync(&s, st->scal, sizeof(s), hipMemcpyDeviceToHost, st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    const int nn = s.nnode;
    if (nn <= 0 || nn > st->node_cap) {
        nbh_set_error("no valid tree (nnode %d)", nn);
        return NBHIP_ERROR;
    }
    return nbh_tree_segment(st, 0, nn, 0, kind, ref, skip, level, pos, mass, rcrit2, quad);
}

extern "C" unsigned nbhip_batch_tree_download(nbhip_state* st, int sim, int* kind, int* ref,
                                              int* skip, int* level, double* pos, double* mass,
                                              double* rcrit2, double* quad)
{
    if (!st || sim < 0 || sim >= st->nsim) return NBHIP_USER_ERROR;
    if (!st->have_tree) {
        nbh_set_error("no tree: call nbhip_gravity first (and criterion != Exact)");
        return NBHIP_USER_ERROR;
    }
    NBH_CHECK(hipSetDevice(st->device));
    SimTab t;
    DevScalars s;
    NBH_CHECK(hipMemcpyAsync(&t, st->sims + sim, sizeof(t), hipMemcpyDeviceToHost, st->stream));
    NBH_CHECK(hipMemcpyAsync(&s, st->scal, sizeof(s), hipMemcpyDeviceToHost, st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    if (t.root < 0) return NBHIP_SUCCESS;            /* no massive body: no tree */
    if (t.end <= t.root || t.end > s.nnode || s.nnode > st->node_cap) {
        nbh_set_error("no valid tree segment for simulation %d", sim);
        return NBHIP_ERROR;
    }
    return nbh_tree_segment(st, t.root, t.end - t.root, st->h_body_off[sim], kind, ref, skip,
                            level, pos, mass, rcrit2, quad);
}

extern "C" unsigned nbhip_direct_accel(nbhip_state* st, int nsample, const int* idx, double* acc)
{
    if (!st || nsample < 0 || (nsample > 0 && (!idx || !acc))) return NBHIP_USER_ERROR;
    if (nsample == 0) return NBHIP_SUCCESS;
    for (int k = 0; k < nsample; ++k)
        if (idx[k] < 0 || idx[k] >= st->n) {
            nbh_set_error("sample index %d out of range", idx[k]);
            return NBHIP_USER_ERROR;
        }
    NBH_CHECK(hipSetDevice(st->device));
    int* d_idx = nullptr;
    double* d_acc = nullptr;
    NBH_CHECK(hipMalloc(&d_idx, nsample * sizeof(int)));
    hipError_t e = hipMalloc(&d_acc, 3 * (size_t) nsample * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_idx, idx, nsample * sizeof(int), hipMemcpyHostToDevice, st->stream);
    if (e == hipSuccess) {
        nbh_direct(st, nsample, d_idx, d_acc);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(acc, d_acc, 3 * (size_t) nsample * sizeof(double),
                           hipMemcpyDeviceToHost, st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    (void) hipFree(d_idx);
    if (d_acc) (void) hipFree(d_acc);
    if (e != hipSuccess) {
        nbh_set_error("direct sum failed: %s", hipGetErrorString(e));
        return NBHIP_GPU_ERROR;
    }
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_walk_accel(nbhip_state* st, double* acc)
{
    if (!st || !acc) return NBHIP_USER_ERROR;
    if (!st->have_tree) {
        nbh_set_error("no tree: call nbhip_gravity first");
        return NBHIP_USER_ERROR;
    }
    NBH_CHECK(hipSetDevice(st->device));
    const size_t n = (size_t) st->n;
    if (n == 0) return NBHIP_SUCCESS;
    double* t = nullptr;
    NBH_CHECK(hipMalloc(&t, 3 * n * sizeof(double)));
    nbh_walk(st, 0, 0, t, t + n, t + 2 * n, 1, nullptr);
    std::vector<double> h(3 * n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(h.data(), t, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    (void) hipFree(t);
    if (e != hipSuccess) {
        nbh_set_error("walk failed: %s", hipGetErrorString(e));
        return NBHIP_GPU_ERROR;
    }
    for (size_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) acc[3 * i + d] = h[d * n + i];
    return NBHIP_SUCCESS;
}

static unsigned nbh_walk_counters(nbhip_state* st, unsigned long long* counts, int ncount)
{
    if (!st || !counts) return NBHIP_USER_ERROR;
    if (!st->have_tree) {
        nbh_set_error("no tree: call nbhip_gravity first");
        return NBHIP_USER_ERROR;
    }
    NBH_CHECK(hipSetDevice(st->device));
    const size_t n = (size_t) st->n;
    double* t = nullptr;
    NBH_CHECK(hipMalloc(&t, 3 * (n ? n : 1) * sizeof(double)));
    unsigned long long zero[5] = { 0, 0, 0, 0, 0 };
    hipError_t e = hipMemcpyAsync(&st->scal->counters[0], zero, sizeof(zero),
                                  hipMemcpyHostToDevice, st->stream);
    if (e == hipSuccess) {
        nbh_walk(st, 0, 1, t, t + n, t + 2 * n, 1, nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(counts, &st->scal->counters[0], ncount * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    (void) hipFree(t);
    if (e != hipSuccess) {
        nbh_set_error("counting walk failed: %s", hipGetErrorString(e));
        return NBHIP_GPU_ERROR;
    }
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_walk_counters(nbhip_state* st, unsigned long long* counts4)
{
    return nbh_walk_counters(st, counts4, 4);
}

extern "C" unsigned nbhip_walk_counters5(nbhip_state* st, unsigned long long* counts5)
{
    return nbh_walk_counters(st, counts5, 5);
}

extern "C" unsigned nbhip_set_profiling(nbhip_state* st, int enable)
{
    if (!st) return NBHIP_USER_ERROR;
    st->profiling = enable ? 1 : 0;
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_phase_times(nbhip_state* st, double* ms5)
{
    if (!st || !ms5) return NBHIP_USER_ERROR;
    for (int k = 0; k < 5; ++k) ms5[k] = st->phase_ms[k];
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_set_tune(nbhip_state* st, const nbhip_tune* tune)
{
    if (!st || !tune) return NBHIP_USER_ERROR;
    if (tune->block != 64 && tune->block != 128 && tune->block != 256) return NBHIP_USER_ERROR;
    if (tune->batch < 1 || (tune->batch & (tune->batch - 1))) {
        nbh_set_error("nbhip_set_tune: batch must be a power of two");
        return NBHIP_USER_ERROR;
    }
    st->tune.block = tune->block;
    st->tune.batch = tune->batch;
    st->graph_ready = 0;
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_set_open_scale(nbhip_state* st, double scale)
{
    if (!st) return NBHIP_USER_ERROR;
    if (st->ctx.precision != NBHIP_PRECISION_MIXED_COHERENT &&
        st->ctx.precision != NBHIP_PRECISION_MIXED_GROUP) {
        nbh_set_error("nbhip_set_open_scale: only the coherent and group mixed walks open more");
        return NBHIP_USER_ERROR;
    }
    if (!(scale >= 1.0) || !(scale <= 1.0e6)) {
        nbh_set_error("nbhip_set_open_scale: scale %.17g is outside [1, 1e6]", scale);
        return NBHIP_USER_ERROR;
    }
    st->open_scale = scale;
    st->graph_ready = 0;          /* the captured step carries the old factor */
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_set_group_expansion(nbhip_state* st, double ratio)
{
    if (!st) return NBHIP_USER_ERROR;
    if (st->ctx.precision != NBHIP_PRECISION_MIXED_GROUP) {
        nbh_set_error("nbhip_set_group_expansion: only the group walk expands far cells");
        return NBHIP_USER_ERROR;
    }
    if (!(ratio >= 0.0) || !(ratio <= 0.5)) {
        nbh_set_error("nbhip_set_group_expansion: ratio %.17g is outside [0, 0.5]", ratio);
        return NBHIP_USER_ERROR;
    }
    st->group_far = ratio;
    st->graph_ready = 0;          /* the captured step carries the old ratio */
    return NBHIP_SUCCESS;
}

/* Project-written autotuner: each candidate walk setting runs three times
 * on the current tree (outputs to scratch), timed with events on the
 * library's stream; the median decides. */
extern "C" unsigned nbhip_autotune(nbhip_state* st, nbhip_tune* best, nbhip_tune* trials,
                                   int max_trials, int* ntrials)
{
    if (!st) return NBHIP_USER_ERROR;
    if (!st->have_tree) {
        nbh_set_error("no tree: call nbhip_gravity first");
        return NBHIP_USER_ERROR;
    }
    NBH_CHECK(hipSetDevice(st->device));
    const size_t n = (size_t) (st->n > 0 ? st->n : 1);
    double* t = nullptr;
    NBH_CHECK(hipMalloc(&t, 3 * n * sizeof(double)));
    const int blocks[3] = { 64, 128, 256 };
    const bool group = st->ctx.precision == NBHIP_PRECISION_MIXED_GROUP;
    /* mixed walks: FP32 partial-sum node block; group walk: LDS list depth */
    const int batches_mixed[4] = { 256, 1024, 4096, 16384 };
    const int depths_group[3] = { 64, 128, 256 };
    const bool mixed = st->ctx.precision != NBHIP_PRECISION_FP64;
    const int nb = group ? 3 : (mixed ? 4 : 1);
    nbhip_tune bestt = { st->tune.block, st->tune.batch, 1e300 };
    int cnt = 0;
    hipError_t e = hipSuccess;
    for (int bi = 0; bi < 3 && e == hipSuccess; ++bi) {
        for (int ci = 0; ci < nb && e == hipSuccess; ++ci) {
            WalkTune w = { blocks[bi], group ? depths_group[ci]
                                           : (mixed ? batches_mixed[ci] : st->tune.batch) };
            if (group && blocks[bi] == 256 && w.batch > 128) continue;   /* LDS limit */
            float ms[3];
            for (int r = 0; r < 3 && e == hipSuccess; ++r) {
                e = hipEventRecord(st->ev[0], st->stream);
                nbh_walk(st, 0, 0, t, t + n, t + 2 * n, 0, &w);
                if (e == hipSuccess) e = hipEventRecord(st->ev[1], st->stream);
                if (e == hipSuccess) e = hipEventSynchronize(st->ev[1]);
                if (e == hipSuccess) e = hipEventElapsedTime(&ms[r], st->ev[0], st->ev[1]);
            }
            if (e != hipSuccess) break;
            float lo = fminf(ms[0], fminf(ms[1], ms[2])), hi = fmaxf(ms[0], fmaxf(ms[1], ms[2]));
            const double med = (double) (ms[0] + ms[1] + ms[2] - lo - hi);
            if (trials && cnt < max_trials) {
                trials[cnt].block = w.block;
                trials[cnt].batch = w.batch;
                trials[cnt].ms = med;
            }
            ++cnt;
            if (med < bestt.ms) {
                bestt.block = w.block;
                bestt.batch = w.batch;
                bestt.ms = med;
            }
        }
    }
    (void) hipFree(t);
    if (e != hipSuccess) {
        nbh_set_error("autotune failed: %s", hipGetErrorString(e));
        return NBHIP_GPU_ERROR;
    }
    st->tune.block = bestt.block;
    st->tune.batch = bestt.batch;
    st->graph_ready = 0;
    if (best) *best = bestt;
    if (ntrials) *ntrials = cnt;
    return NBHIP_SUCCESS;
}

/* ---------------------------------------------------------------------- */
/* Adjoint leapfrog with checkpointing                                    */
/* ---------------------------------------------------------------------- */
/* state slot <-> live arrays: x y z vx vy vz ax ay az, then the SimTab rows
 * (root sizes persist across steps, so a recomputed step needs them) */
static hipError_t nbh_slot_save(nbhip_state* st, double* slot, SimTab* tab)
{
    const size_t N = (size_t) st->n;
    double* src[9] = { st->px, st->py, st->pz, st->vx, st->vy, st->vz, st->ax, st->ay, st->az };
    hipError_t e = hipSuccess;
    for (int k = 0; k < 9 && e == hipSuccess && N; ++k)
        e = hipMemcpyAsync(slot + k * N, src[k], N * sizeof(double), hipMemcpyDeviceToDevice,
                           st->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(tab, st->sims, st->nsim * sizeof(SimTab), hipMemcpyDeviceToDevice,
                           st->stream);
    return e;
}

static hipError_t nbh_slot_load(nbhip_state* st, const double* slot, const SimTab* tab,
                                int positions_only)
{
    const size_t N = (size_t) st->n;
    double* dst[9] = { st->px, st->py, st->pz, st->vx, st->vy, st->vz, st->ax, st->ay, st->az };
    hipError_t e = hipSuccess;
    const int cnt = positions_only ? 3 : 9;
    for (int k = 0; k < cnt && e == hipSuccess && N; ++k)
        e = hipMemcpyAsync(dst[k], slot + k * N, N * sizeof(double), hipMemcpyDeviceToDevice,
                           st->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(st->sims, tab, st->nsim * sizeof(SimTab), hipMemcpyDeviceToDevice,
                           st->stream);
    return e;
}

/* rebuilds the tree for the positions now in px, py, pz (nothing for Exact) */
static void nbh_adj_rebuild(nbhip_state* st)
{
    if (st->ctx.criterion == NBHIP_EXACT) return;
    nbh_bbox(st, 0);
    nbh_build(st, 0);
}

/* adjoint vectors and node records only (no checkpoints): for nbhip_gravity_vjp */
static unsigned nbh_adj_vectors(nbhip_state* st)
{
    if (st->adj && st->adj->g) return NBHIP_SUCCESS;
    if (!st->adj) st->adj = (NbhAdjoint*) calloc(1, sizeof(NbhAdjoint));
    NbhAdjoint* A = st->adj;
    const size_t N = (size_t) st->n, ns = (size_t) st->nsim;
    hipError_t e = nbh_alloc(&A->g, 10 * N);
    if (e == hipSuccess) e = nbh_alloc(&A->gdt, ns);
    if (e == hipSuccess) e = nbh_alloc(&A->gnode, 10 * (size_t) st->node_cap);
    if (e != hipSuccess) {
        nbh_set_error("adjoint vectors: %s", hipGetErrorString(e));
        nbh_adj_free(st);
        return NBHIP_GPU_ERROR;
    }
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_gravity_vjp(nbhip_state* st, const double* gacc, double* gpos,
                                      double* gmass)
{
    if (!st || !gacc) return NBHIP_USER_ERROR;
    if (st->ctx.precision != NBHIP_PRECISION_FP64) {
        nbh_set_error("nbhip_gravity_vjp: needs an FP64 state");
        return NBHIP_UNSUPPORTED;
    }
    if (st->ctx.criterion != NBHIP_EXACT && !st->have_tree) {
        nbh_set_error("nbhip_gravity_vjp: no tree; call nbhip_gravity first");
        return NBHIP_USER_ERROR;
    }
    NBH_CHECK(hipSetDevice(st->device));
    unsigned rc = nbh_adj_vectors(st);
    if (rc) return rc;
    NbhAdjoint* A = st->adj;
    const size_t N = (size_t) st->n;
    std::vector<double> h(10 * N, 0.0);
    for (size_t i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d) h[(6 + d) * N + i] = gacc[3 * i + d];
    NBH_CHECK(hipMemcpyAsync(A->g, h.data(), 10 * N * sizeof(double), hipMemcpyHostToDevice,
                             st->stream));
    nbh_adj_force(st);
    nbh_adj_body(st, 0, nullptr, nullptr);
    NBH_CHECK(hipGetLastError());
    NBH_CHECK(hipMemcpyAsync(h.data(), A->g, 10 * N * sizeof(double), hipMemcpyDeviceToHost,
                             st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    for (size_t i = 0; i < N; ++i) {
        for (int d = 0; d < 3; ++d)
            if (gpos) gpos[3 * i + d] = h[d * N + i];
        if (gmass) gmass[i] = h[9 * N + i];
    }
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_adjoint_forward(nbhip_state* st, unsigned nsteps, unsigned every)
{
    if (!st) return NBHIP_USER_ERROR;
    if (st->ctx.precision != NBHIP_PRECISION_FP64) {
        nbh_set_error("nbhip_adjoint_forward: the adjoint needs an FP64 state");
        return NBHIP_UNSUPPORTED;
    }
    if (every < 1 || nsteps < 1) {
        nbh_set_error("nbhip_adjoint_forward: nsteps and every must be >= 1");
        return NBHIP_USER_ERROR;
    }
    if (every > nsteps) every = nsteps;
    NBH_CHECK(hipSetDevice(st->device));
    nbh_adj_free(st);
    NbhAdjoint* A = (NbhAdjoint*) calloc(1, sizeof(NbhAdjoint));
    st->adj = A;
    A->nsteps = nsteps;
    A->every = every;
    A->nckpt = nsteps / every + 1;
    const size_t N = (size_t) st->n, ns = (size_t) st->nsim;
    hipError_t e = hipSuccess;
#define A_(ptr, cnt) if (e == hipSuccess) e = nbh_alloc(&A->ptr, (cnt))
    A_(ckpt, 9 * N * A->nckpt);
    A_(ckpt_tab, ns * A->nckpt);
    A_(seg, 9 * N * (every + 1));
    A_(seg_tab, ns * (every + 1));
    A_(fin, 9 * N);
    A_(fin_tab, ns);
    A_(g, 10 * N);
    A_(gdt, ns);
    A_(gnode, 10 * (size_t) st->node_cap);
#undef A_
    if (e != hipSuccess) {
        nbh_set_error("adjoint storage for %u checkpoints failed: %s", A->nckpt,
                      hipGetErrorString(e));
        nbh_adj_free(st);
        return NBHIP_GPU_ERROR;
    }
    unsigned rc = nbhip_gravity(st);
    if (NBHIP_STATUS_IS_FATAL(rc)) return rc;
    NBH_CHECK(nbh_slot_save(st, A->ckpt, A->ckpt_tab));
    for (unsigned k = 1; k <= nsteps; ++k) {
        rc |= nbhip_step(st);
        if (NBHIP_STATUS_IS_FATAL(rc)) return rc;
        if (k % every == 0)
            NBH_CHECK(nbh_slot_save(st, A->ckpt + 9 * N * (k / every),
                                    A->ckpt_tab + ns * (k / every)));
    }
    NBH_CHECK(hipStreamSynchronize(st->stream));
    A->ready = 1;
    return rc;
}

extern "C" unsigned nbhip_adjoint_backward(nbhip_state* st, const double* gpos, const double* gvel,
                                           double* gpos0, double* gvel0, double* gmass,
                                           double* gdt)
{
    if (!st || !gpos || !gvel) return NBHIP_USER_ERROR;
    NbhAdjoint* A = st->adj;
    if (!A || !A->ready) {
        nbh_set_error("nbhip_adjoint_backward: run nbhip_adjoint_forward first");
        return NBHIP_USER_ERROR;
    }
    NBH_CHECK(hipSetDevice(st->device));
    const size_t N = (size_t) st->n, ns = (size_t) st->nsim;
    const unsigned every = A->every;
    const double dt = st->ctx.timestep;
    /* adjoint of the final state: AoS host -> SoA device */
    std::vector<double> h(10 * N, 0.0);
    for (size_t i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d) {
            h[d * N + i] = gpos[3 * i + d];
            h[(3 + d) * N + i] = gvel[3 * i + d];
        }
    NBH_CHECK(hipMemcpyAsync(A->g, h.data(), 10 * N * sizeof(double), hipMemcpyHostToDevice,
                             st->stream));
    NBH_CHECK(hipMemsetAsync(A->gdt, 0, ns * sizeof(double), st->stream));
    NBH_CHECK(nbh_slot_save(st, A->fin, A->fin_tab));
    const unsigned step_fin = st->step;
    const int prof = st->profiling;
    st->profiling = 0;

    unsigned rc = 0;
    const int nseg = (int) ((A->nsteps + every - 1) / every);
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const unsigned s0 = (unsigned) sg * every;
        const unsigned len = (A->nsteps - s0 < every) ? A->nsteps - s0 : every;
        /* recompute the segment from its checkpoint, keeping every state */
        NBH_CHECK(nbh_slot_load(st, A->ckpt + 9 * N * sg, A->ckpt_tab + ns * sg, 0));
        NBH_CHECK(nbh_slot_save(st, A->seg, A->seg_tab));
        for (unsigned k = 1; k <= len; ++k) {
            rc |= nbh_force(st, 1, 1);
            NBH_CHECK(nbh_slot_save(st, A->seg + 9 * N * k, A->seg_tab + ns * k));
        }
        rc |= nbh_collect(st);
        if (NBHIP_STATUS_IS_FATAL(rc)) break;
        /* reverse through the segment: state k = x_{n+1}, state k-1 = x_n */
        for (unsigned k = len; k >= 1; --k) {
            const double* cur = A->seg + 9 * N * k;
            const double* prev = A->seg + 9 * N * (k - 1);
            nbh_adj_kick_end(st, cur + 6 * N);
            NBH_CHECK(nbh_slot_load(st, cur, A->seg_tab + ns * k, 1));
            nbh_adj_rebuild(st);
            nbh_adj_force(st);
            nbh_adj_body(st, 1, prev + 3 * N, prev + 6 * N);
        }
    }
    if (!NBHIP_STATUS_IS_FATAL(rc)) {
        /* the force evaluation before the first step */
        NBH_CHECK(nbh_slot_load(st, A->ckpt, A->ckpt_tab, 1));
        nbh_adj_rebuild(st);
        nbh_adj_force(st);
        nbh_adj_body(st, 0, nullptr, nullptr);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            nbh_set_error("adjoint launch failed: %s", hipGetErrorString(e));
            rc |= NBHIP_GPU_ERROR;
        }
    }
    /* the state goes back to the end of the forward run, tree included */
    NBH_CHECK(nbh_slot_load(st, A->fin, A->fin_tab, 0));
    nbh_adj_rebuild(st);
    st->step = step_fin;
    st->profiling = prof;
    rc |= nbh_collect(st);
    if (NBHIP_STATUS_IS_FATAL(rc)) return rc;

    NBH_CHECK(hipMemcpyAsync(h.data(), A->g, 10 * N * sizeof(double), hipMemcpyDeviceToHost,
                             st->stream));
    std::vector<double> hdt(ns);
    NBH_CHECK(hipMemcpyAsync(hdt.data(), A->gdt, ns * sizeof(double), hipMemcpyDeviceToHost,
                             st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    for (size_t i = 0; i < N; ++i)
        for (int d = 0; d < 3; ++d) {
            if (gpos0) gpos0[3 * i + d] = h[d * N + i];
            if (gvel0) gvel0[3 * i + d] = h[(3 + d) * N + i];
        }
    if (gmass)
        for (size_t i = 0; i < N; ++i) gmass[i] = h[9 * N + i];
    if (gdt)
        for (size_t k = 0; k < ns; ++k) gdt[k] = hdt[k];
    return rc;
}

/* ---------------------------------------------------------------------- */
/* Histogram                                                              */
/* ---------------------------------------------------------------------- */
extern "C" unsigned nbhip_set_ignore(nbhip_state* st, const unsigned char* ignore)
{
    if (!st) return NBHIP_USER_ERROR;
    NBH_CHECK(hipSetDevice(st->device));
    if (!ignore) {
        if (st->ignore) NBH_CHECK(hipFree(st->ignore));
        st->ignore = nullptr;
        return NBHIP_SUCCESS;
    }
    if (!st->ignore) NBH_CHECK(nbh_alloc(&st->ignore, (size_t) st->n));
    if (st->n)
        NBH_CHECK(hipMemcpyAsync(st->ignore, ignore, (size_t) st->n, hipMemcpyHostToDevice,
                                 st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_histogram(nbhip_state* st, const nbhip_histogram_params* hp,
                                    unsigned* counts, unsigned* totals)
{
    if (!st || !hp) return NBHIP_USER_ERROR;
    if (hp->lambda_bins < 1 || hp->beta_bins < 1 || !(hp->lambda_end > hp->lambda_start) ||
        !(hp->beta_end > hp->beta_start)) {
        nbh_set_error("nbhip_histogram: empty bin range");
        return NBHIP_USER_ERROR;
    }
    NBH_CHECK(hipSetDevice(st->device));
    const size_t nbin = (size_t) hp->lambda_bins * (size_t) hp->beta_bins;
    const size_t need = (size_t) st->nsim * (nbin + 1);
    if (need > st->hist_cap) {
        if (st->hist_buf) NBH_CHECK(hipFree(st->hist_buf));
        st->hist_buf = nullptr;
        st->hist_cap = 0;
        NBH_CHECK(nbh_alloc(&st->hist_buf, need));
        st->hist_cap = need;
    }
    unsigned* d_counts = st->hist_buf;
    unsigned* d_totals = st->hist_buf + (size_t) st->nsim * nbin;
    NBH_CHECK(hipMemsetAsync(st->hist_buf, 0, need * sizeof(unsigned), st->stream));
    nbh_histogram(st, hp, d_counts, d_totals);
    NBH_CHECK(hipGetLastError());
    if (counts)
        NBH_CHECK(hipMemcpyAsync(counts, d_counts, (size_t) st->nsim * nbin * sizeof(unsigned),
                                 hipMemcpyDeviceToHost, st->stream));
    if (totals)
        NBH_CHECK(hipMemcpyAsync(totals, d_totals, (size_t) st->nsim * sizeof(unsigned),
                                 hipMemcpyDeviceToHost, st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_histogram_dispersion(nbhip_state* st, int sim,
                                               const nbhip_histogram_params* hp,
                                               const nbhip_disp_params* dp, double* out)
{
    if (!st || !hp || !dp || !out || sim < 0 || sim >= st->nsim) return NBHIP_USER_ERROR;
    if (hp->lambda_bins < 1 || hp->beta_bins < 1 || !(hp->lambda_end > hp->lambda_start) ||
        !(hp->beta_end > hp->beta_start) || dp->iter_max < 0 || dp->iter_max > 1000) {
        nbh_set_error("nbhip_histogram_dispersion: bad histogram or rejection settings");
        return NBHIP_USER_ERROR;
    }
    NBH_CHECK(hipSetDevice(st->device));
    const int off = st->h_body_off[sim], cnt = st->h_body_n[sim];
    const size_t nbin = (size_t) hp->lambda_bins * (size_t) hp->beta_bins;
    if (cnt <= 0) {
        /* no bodies: every dispersion 0, every average 0 with error -1 */
        for (size_t k = 0; k < 14; ++k)
            for (size_t b = 0; b < nbin; ++b) out[k * nbin + b] = (k >= 4 && k % 2) ? -1.0 : 0.0;
        return NBHIP_SUCCESS;
    }
    const size_t nblk = ((size_t) cnt + 1023) / 1024;
    /* one scratch block, kept by the state and grown on demand: body bins,
     * five values per body, per-workgroup partial sums, per-bin state and
     * the 14 output rows, each section 256-byte aligned */
    auto up = [](size_t bytes) { return (bytes + 255) & ~(size_t) 255; };
    const size_t o_val = up((size_t) cnt * sizeof(int));
    const size_t o_part = o_val + up(5 * (size_t) cnt * sizeof(double));
    const size_t n_part = 16 * nblk * nbin;
    if (n_part > ((size_t) 1 << 31)) {
        nbh_set_error("nbhip_histogram_dispersion: %zu bins x %d bodies is too large", nbin, cnt);
        return NBHIP_USER_ERROR;
    }
    const size_t o_state = o_part + up(n_part * sizeof(double));
    const size_t o_out = o_state + up(32 * nbin * sizeof(double));
    const size_t need = o_out + up(14 * nbin * sizeof(double));
    hipError_t e = hipSuccess;
    if (need > st->disp_cap) {
        if (st->disp_buf) e = hipFree(st->disp_buf);
        st->disp_buf = nullptr;
        st->disp_cap = 0;
        if (e == hipSuccess) e = nbh_alloc(&st->disp_buf, need);
        if (e == hipSuccess) st->disp_cap = need;
    }
    if (e == hipSuccess) {
        char* b = st->disp_buf;
        double* dout = reinterpret_cast<double*>(b + o_out);
        nbh_dispersion(st, off, cnt, hp, dp, reinterpret_cast<int*>(b),
                       reinterpret_cast<double*>(b + o_val), reinterpret_cast<double*>(b + o_part),
                       reinterpret_cast<double*>(b + o_state), dout);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(out, dout, 14 * nbin * sizeof(double), hipMemcpyDeviceToHost,
                               st->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    if (e != hipSuccess) {
        nbh_set_error("nbhip_histogram_dispersion: %s", hipGetErrorString(e));
        return NBHIP_GPU_ERROR;
    }
    return NBHIP_SUCCESS;
}

/* Reverse orbit with its t_rev tangent (forward mode): each KDK step is
 * differentiated alongside, d(dt)/d(t_rev) = 1 / nsteps, and the change of
 * the acceleration is the potential's Jacobian times the position tangent
 * (nbh_ext_jac, symmetric). */
extern "C" unsigned nbhip_orbit_tangent(const nbhip_potential* pot, const double* pos,
                                        const double* vel, double t_rev, int nsteps,
                                        double* out_pos, double* out_vel, double* dpos,
                                        double* dvel)
{
    if (!pot || !pos || !vel || !out_pos || !out_vel || nsteps < 1 || !std::isfinite(t_rev)) {
        nbh_set_error("nbhip_orbit_tangent: bad arguments");
        return NBHIP_USER_ERROR;
    }
    const double dt = t_rev / nsteps;
    const double half = 0.5 * dt;
    const double ddt = 1.0 / nsteps, dhalf = 0.5 * ddt;
    double x[3] = { pos[0], pos[1], pos[2] };
    double v[3] = { -vel[0], -vel[1], -vel[2] };
    double X[3] = { 0.0, 0.0, 0.0 }, V[3] = { 0.0, 0.0, 0.0 }, A[3] = { 0.0, 0.0, 0.0 };
    double a[3];
    nbh_ext_accel(*pot, x[0], x[1], x[2], a[0], a[1], a[2]);
    for (int k = 0; k < nsteps; ++k) {
        for (int d = 0; d < 3; ++d) {
            V[d] = V[d] + (dhalf * a[d] + half * A[d]);
            v[d] = v[d] + half * a[d];
        }
        for (int d = 0; d < 3; ++d) {
            X[d] = X[d] + (ddt * v[d] + dt * V[d]);
            x[d] = x[d] + dt * v[d];
        }
        nbh_ext_accel(*pot, x[0], x[1], x[2], a[0], a[1], a[2]);
        nbh_ext_jac(*pot, x[0], x[1], x[2], X[0], X[1], X[2], A[0], A[1], A[2]);
        for (int d = 0; d < 3; ++d) {
            V[d] = V[d] + (dhalf * a[d] + half * A[d]);
            v[d] = v[d] + half * a[d];
        }
    }
    for (int d = 0; d < 3; ++d) {
        out_pos[d] = x[d];
        out_vel[d] = -v[d];
        if (dpos) dpos[d] = X[d];
        if (dvel) dvel[d] = -V[d];
    }
    return NBHIP_SUCCESS;
}
