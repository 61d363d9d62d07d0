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
    if (c->precision != NBHIP_PRECISION_FP64 && c->precision != NBHIP_PRECISION_MIXED) {
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

static void nbh_free_all(nbhip_state* st)
{
    void* ptrs[] = { st->px,   st->py,     st->pz,    st->vx,      st->vy,     st->vz,
                     st->m,    st->ax,     st->ay,    st->az,      st->key,    st->key_alt,
                     st->perm, st->perm_alt, st->hist, st->dtot, st->spos,   st->su,     st->dadj,
                     st->cincl, st->scan_part, st->blo, st->bhi,   st->bdelta, st->bparent,
                     st->node, st->nquad,  st->nmid,  st->nlevel,  st->nparent,
                     st->level_list, st->nodef, st->scal };
    for (void* p : ptrs)
        if (p) (void) hipFree(p);
}

extern "C" unsigned nbhip_create(const nbhip_ctx* ctx, int device, int nbody, nbhip_state** out)
{
    if (!ctx || !out || nbody < 0) {
        nbh_set_error("nbhip_create: bad arguments");
        return NBHIP_USER_ERROR;
    }
    *out = nullptr;
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

    nbhip_state* st = (nbhip_state*) calloc(1, sizeof(nbhip_state));
    st->ctx = *ctx;
    st->device = device;
    st->n = nbody;
    st->node_cap = 5 * nbody + 1024;
    st->sort_blocks = nbody > 0 ? (nbody + NBH_SORT_TILE - 1) / NBH_SORT_TILE : 1;
    const size_t n = (size_t) nbody;
    const size_t nc = (size_t) st->node_cap;
    hipError_t e = hipSuccess;
#define A(ptr, cnt) if (e == hipSuccess) e = nbh_alloc(&st->ptr, (cnt))
    A(px, n); A(py, n); A(pz, n); A(vx, n); A(vy, n); A(vz, n); A(m, n);
    A(ax, n); A(ay, n); A(az, n);
    A(key, n); A(key_alt, n); A(perm, n); A(perm_alt, n);
    A(hist, 256 * (size_t) st->sort_blocks);
    A(dtot, 256);
    A(spos, n); A(su, 3 * n); A(dadj, n); A(cincl, n);
    A(scan_part, 4096);
    A(blo, n); A(bhi, n); A(bdelta, n); A(bparent, 2 * n);
    A(node, nc); A(nquad, 6 * nc); A(nmid, 3 * nc); A(nlevel, nc); A(nparent, nc);
    A(level_list, nc);
    if (ctx->precision == NBHIP_PRECISION_MIXED) A(nodef, nc);
    A(scal, 1);
#undef A
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking);
    for (int k = 0; k < 6 && e == hipSuccess; ++k) e = hipEventCreate(&st->ev[k]);
    if (e != hipSuccess) {
        nbh_set_error("allocation for %d bodies failed: %s", nbody, hipGetErrorString(e));
        nbh_free_all(st);
        free(st);
        return NBHIP_GPU_ERROR;
    }
    DevScalars s0;
    memset(&s0, 0, sizeof(s0));
    s0.rsize = ctx->tree_rsize;
    s0.key_scale = 9223372036854775808.0 / ctx->tree_rsize;
    NBH_CHECK(hipMemcpy(st->scal, &s0, sizeof(s0), hipMemcpyHostToDevice));
    NBH_CHECK(hipMemset(st->ax, 0, n * sizeof(double) + (n ? 0 : 8)));
    NBH_CHECK(hipMemset(st->ay, 0, n * sizeof(double) + (n ? 0 : 8)));
    NBH_CHECK(hipMemset(st->az, 0, n * sizeof(double) + (n ? 0 : 8)));
    *out = st;
    return NBHIP_SUCCESS;
}

extern "C" void nbhip_release(nbhip_state* st)
{
    if (!st) return;
    (void) hipSetDevice(st->device);
    if (st->stream) (void) hipStreamSynchronize(st->stream);
    for (int k = 0; k < 6; ++k)
        if (st->ev[k]) (void) hipEventDestroy(st->ev[k]);
    if (st->stream) (void) hipStreamDestroy(st->stream);
    nbh_free_all(st);
    free(st);
}

extern "C" unsigned nbhip_upload(nbhip_state* st, const double* pos, const double* vel,
                                 const double* mass)
{
    if (!st || (st->n > 0 && (!pos || !vel || !mass))) {
        nbh_set_error("nbhip_upload: bad arguments");
        return NBHIP_USER_ERROR;
    }
    NBH_CHECK(hipSetDevice(st->device));
    const size_t n = (size_t) st->n;
    if (n == 0) return NBHIP_SUCCESS;
    std::vector<double> tmp(n);
    double* dst[6] = { st->px, st->py, st->pz, st->vx, st->vy, st->vz };
    for (int k = 0; k < 6; ++k) {
        const double* src = k < 3 ? pos : vel;
        const int d = k % 3;
        for (size_t i = 0; i < n; ++i) tmp[i] = src[3 * i + d];
        NBH_CHECK(hipMemcpyAsync(dst[k], tmp.data(), n * sizeof(double), hipMemcpyHostToDevice,
                                 st->stream));
        NBH_CHECK(hipStreamSynchronize(st->stream));
    }
    NBH_CHECK(hipMemcpyAsync(st->m, mass, n * sizeof(double), hipMemcpyHostToDevice, st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    st->have_tree = 0;
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_download(nbhip_state* st, double* pos, double* vel, double* acc)
{
    if (!st) {
        nbh_set_error("nbhip_download: null state");
        return NBHIP_USER_ERROR;
    }
    NBH_CHECK(hipSetDevice(st->device));
    const size_t n = (size_t) st->n;
    if (n == 0) return NBHIP_SUCCESS;
    std::vector<double> tmp(n);
    double* outs[3] = { pos, vel, acc };
    double* srcs[3][3] = { { st->px, st->py, st->pz },
                           { st->vx, st->vy, st->vz },
                           { st->ax, st->ay, st->az } };
    for (int o = 0; o < 3; ++o) {
        if (!outs[o]) continue;
        for (int d = 0; d < 3; ++d) {
            NBH_CHECK(hipMemcpyAsync(tmp.data(), srcs[o][d], n * sizeof(double),
                                     hipMemcpyDeviceToHost, st->stream));
            NBH_CHECK(hipStreamSynchronize(st->stream));
            for (size_t i = 0; i < n; ++i) outs[o][3 * i + d] = tmp[i];
        }
    }
    return NBHIP_SUCCESS;
}

/* Read back the device checks and fold them into the sticky status,
 * mapping a missed self-interaction onto the incest codes
 * (nbIncestStatusCheck, nbody/src/nbody_grav.c:295-303). */
static unsigned nbh_collect(nbhip_state* st)
{
    DevScalars s;
    NBH_CHECK(hipMemcpyAsync(&s, st->scal, sizeof(s), hipMemcpyDeviceToHost, st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    unsigned rc = st->status | s.status;
    if (s.incest)
        rc |= st->ctx.allow_incest ? NBHIP_TREE_INCEST_NONFATAL : NBHIP_TREE_INCEST_FATAL;
    st->status = rc;
    if (rc & NBHIP_CELL_OVERFLOW_ERROR)
        nbh_set_error("node table overflow: %d nodes for capacity %d", s.nnode, st->node_cap);
    else if (rc & NBHIP_TREE_STRUCTURE_ERROR)
        nbh_set_error("tree structure error (max depth %d, rsize %.17g)", s.maxdepth, s.rsize);
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
        nbh_build(st);
    }
    nbh_walk(st, kick, 0, st->ax, st->ay, st->az, 0);
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
    st->step++;
    rc = nbh_collect(st);
    nbh_phase_read(st);
    return rc;
}

extern "C" unsigned nbhip_steps(nbhip_state* st, unsigned nsteps)
{
    if (!st) return NBHIP_USER_ERROR;
    unsigned rc = st->status;
    if (NBHIP_STATUS_IS_FATAL(rc)) return rc;
    for (unsigned k = 0; k < nsteps; ++k) {
        rc |= nbhip_step(st);
        if (NBHIP_STATUS_IS_FATAL(rc)) return rc;
    }
    return rc;
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

extern "C" unsigned nbhip_tree_info_get(nbhip_state* st, nbhip_tree_info* info)
{
    if (!st || !info) return NBHIP_USER_ERROR;
    NBH_CHECK(hipSetDevice(st->device));
    DevScalars s;
    NBH_CHECK(hipMemcpyAsync(&s, st->scal, sizeof(s), hipMemcpyDeviceToHost, st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    info->rsize = s.rsize;
    info->ncell = s.ncell;
    info->max_depth = s.maxdepth;
    info->nnode = s.nnode;
    info->nmassive = s.nmassive;
    info->step = st->step;
    info->status = st->status;
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
    std::vector<NodeW> nodes(nn);
    std::vector<double> q(6 * (size_t) nn);
    std::vector<int> lev(nn);
    std::vector<unsigned> perm(st->n > 0 ? st->n : 1);
    NBH_CHECK(hipMemcpy(nodes.data(), st->node, nn * sizeof(NodeW), hipMemcpyDeviceToHost));
    NBH_CHECK(hipMemcpy(q.data(), st->nquad, 6 * (size_t) nn * sizeof(double),
                        hipMemcpyDeviceToHost));
    NBH_CHECK(hipMemcpy(lev.data(), st->nlevel, nn * sizeof(int), hipMemcpyDeviceToHost));
    if (st->n > 0)
        NBH_CHECK(hipMemcpy(perm.data(), st->perm, st->n * sizeof(unsigned),
                            hipMemcpyDeviceToHost));
    for (int i = 0; i < nn; ++i) {
        const NodeW& w = nodes[i];
        const bool body = w.body >= 0;
        if (kind) kind[i] = body ? 1 : 0;
        if (ref) ref[i] = body ? (int) perm[w.body] : -1;
        if (skip) skip[i] = w.next;
        if (level) level[i] = body ? -1 : lev[i];
        if (pos) { pos[3 * i] = w.x; pos[3 * i + 1] = w.y; pos[3 * i + 2] = w.z; }
        if (mass) mass[i] = w.m;
        if (rcrit2) rcrit2[i] = body ? 0.0 : w.rcrit2;
        if (quad)
            for (int k = 0; k < 6; ++k) quad[6 * i + k] = body ? 0.0 : q[6 * (size_t) i + k];
    }
    return NBHIP_SUCCESS;
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
    nbh_walk(st, 0, 0, t, t + n, t + 2 * n, 1);
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

extern "C" unsigned nbhip_walk_counters(nbhip_state* st, unsigned long long* counts4)
{
    if (!st || !counts4) return NBHIP_USER_ERROR;
    if (!st->have_tree) {
        nbh_set_error("no tree: call nbhip_gravity first");
        return NBHIP_USER_ERROR;
    }
    NBH_CHECK(hipSetDevice(st->device));
    const size_t n = (size_t) st->n;
    double* t = nullptr;
    NBH_CHECK(hipMalloc(&t, 3 * (n ? n : 1) * sizeof(double)));
    unsigned long long zero[4] = { 0, 0, 0, 0 };
    hipError_t e = hipMemcpyAsync(&st->scal->counters[0], zero, sizeof(zero),
                                  hipMemcpyHostToDevice, st->stream);
    if (e == hipSuccess) {
        nbh_walk(st, 0, 1, t, t + n, t + 2 * n, 1);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(counts4, &st->scal->counters[0], 4 * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, st->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st->stream);
    (void) hipFree(t);
    if (e != hipSuccess) {
        nbh_set_error("counting walk failed: %s", hipGetErrorString(e));
        return NBHIP_GPU_ERROR;
    }
    return NBHIP_SUCCESS;
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
