/*
 * Lambda-beta histogram of the bodies on the GPU: nbCreateHistogram's raw
 * counts (nbody/src/nbody_histogram.c:860-900) per simulation, so a
 * likelihood evaluation reads 4 bytes per bin instead of every body.
 *
 * Each body not flagged as ignored (ignoreBody: the dark matter of a mixed
 * dwarf) is rotated into the stream frame as nbXYZToLambdaBeta does
 * (nbody/src/nbody_coordinates.c:129-164): X shifted by the Sun's
 * Galactocentric distance, rotated by the (phi, theta, psi) matrix, lambda =
 * atan2(y', x'), beta = asin(-z'/r), both in degrees; its bin is
 * floor((value - start) / size) and bodies outside the range are not
 * counted.  A block whose bodies all belong to one simulation counts in LDS
 * and adds its non-zero bins once; otherwise bodies add to global counters
 * directly.
 */
#include "nbhip_device.h"

#include <algorithm>
#include <mutex>

struct HistParams {
    double r[3][3];         /* rotation rows */
    double sun;
    double lstart, lsize, bstart, bsize;
    int lbins, bbins;
};

#define NBH_HIST_LDS 4096   /* bins a single-simulation block counts in LDS */

/* nbXYZToLambdaBeta + the bin of nbCreateHistogram: returns the bin
 * (lambdaIndex * betaBins + betaIndex) or -1 outside the histogram; beta in
 * degrees. */
__device__ __forceinline__ int nbh_lambda_beta_bin(const HistParams& h, double x, double y,
                                                   double z, double& beta)
{
    const double X = x + h.sun, Y = y, Z = z;
    const double tx = h.r[0][0] * X + h.r[0][1] * Y + h.r[0][2] * Z;
    const double ty = h.r[1][0] * X + h.r[1][1] * Y + h.r[1][2] * Z;
    const double tz = -(h.r[2][0] * X + h.r[2][1] * Y + h.r[2][2] * Z);
    const double r = sqrt(tx * tx + ty * ty + tz * tz);
    const double lambda = atan2(ty, tx) * (180.0 / M_PI);
    beta = asin(tz / r) * (180.0 / M_PI);
    const double li = floor((lambda - h.lstart) / h.lsize);
    const double bi = floor((beta - h.bstart) / h.bsize);
    if (li >= 0.0 && li < (double) h.lbins && bi >= 0.0 && bi < (double) h.bbins)
        return (int) li * h.bbins + (int) bi;
    return -1;
}

__global__ __launch_bounds__(256) void k_histogram(int n, const double* __restrict__ px,
                                                   const double* __restrict__ py,
                                                   const double* __restrict__ pz,
                                                   const unsigned char* __restrict__ ignore,
                                                   const int* __restrict__ simid, HistParams h,
                                                   unsigned* __restrict__ counts,
                                                   unsigned* __restrict__ totals)
{
    __shared__ unsigned lds[NBH_HIST_LDS];
    __shared__ int s_sim;
    __shared__ int s_mixed;
    const int nbin = h.lbins * h.bbins;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < n;
    const int sim = valid ? simid[i] : -1;
    if (threadIdx.x == 0) {
        s_sim = simid[blockIdx.x * 256];
        s_mixed = 0;
    }
    for (int k = threadIdx.x; k < nbin && k < NBH_HIST_LDS; k += 256) lds[k] = 0u;
    __syncthreads();
    if (valid && sim != s_sim) s_mixed = 1;       /* benign race: any writer sets 1 */
    __syncthreads();
    const bool local = !s_mixed && nbin <= NBH_HIST_LDS;

    int bin = -1;
    if (valid && !(ignore && ignore[i])) {
        double beta;
        bin = nbh_lambda_beta_bin(h, px[i], py[i], pz[i], beta);
    }
    if (local) {
        if (bin >= 0) atomicAdd(&lds[bin], 1u);
        __syncthreads();
        unsigned tot = 0;
        for (int k = threadIdx.x; k < nbin; k += 256) {
            const unsigned c = lds[k];
            if (c) atomicAdd(&counts[(size_t) s_sim * nbin + k], c);
            tot += c;
        }
        tot = nbh_wave_sum_u32(tot);
        if (__lane_id() == 0 && tot) atomicAdd(&totals[s_sim], tot);
    } else if (bin >= 0) {
        atomicAdd(&counts[(size_t) sim * nbin + bin], 1u);
        atomicAdd(&totals[sim], 1u);
    }
}

/* ---------------------------------------------------------------------- */
/* per-bin dispersions and averages of beta, v_los, distance, proper motion */
/* ---------------------------------------------------------------------- */
/* nbCreateHistogram's histograms 1-7 (nbody/src/nbody_histogram.c:858-1110):
 * the beta and line-of-sight velocity dispersions, the average line-of-sight
 * velocity, beta, distance and proper motions in declination and right
 * ascension, each after nbCalcDisp and the rounds of nbRemoveOutliers
 * (nbody/src/nbody_mass.c:195-316; calc_vLOS and calc_distance :171-187;
 * nbVXVYVZtomuDec / RA, nbody/src/nbody_coordinates.c:192-287), for the
 * bodies of one simulation.  The averages go through the same rounds as the
 * dispersions of their quantity (histograms 3 and 2 see the same values,
 * sigma and correction; so do 4 and 1), so five quantities suffice: q = 0
 * beta, 1 v_los, 2 distance, 3 mu_dec, 4 mu_ra.
 *
 * k_disp_values records each body's bin (-1: ignored or outside) and the
 * three values once.  Each round of sums is k_disp_part, one workgroup per
 * 1024 consecutive bodies with a thread per bin that adds the workgroup's
 * bodies in body order, then k_disp_finish, a thread per bin that adds the
 * workgroups' partial sums in workgroup order and applies nbCalcDisp.  The
 * sums are therefore deterministic (independent of timing and batching) and
 * differ from the reference's single running sum only in association.
 * Round 0 takes every binned body; each later round keeps a body when
 * |mean - value| < sigma * dispersion of the previous round. */
#define NBH_DISP_TILE 1024
#define NBH_DISP_Q 5
#define NBH_DISP_W (3 * NBH_DISP_Q + 1)   /* per q: sum, sq, removed; then count */
#define NBH_DISP_S 32                     /* per q: sum sq removed var err -; [30] raw count */
#define NBH_DISP_RAW 30

struct DispRound {
    double sigma[NBH_DISP_Q];
    double correct[NBH_DISP_Q];
};

/* The proper-motion frame: the Galactic -> equatorial rotation of
 * cartesianalign (nbody/src/nbody_coordinates.c:192-224) and the Sun's
 * velocity. */
struct PMFrame {
    double rot[3][3];
    double sun_vel[3];
};

/* nbVXVYVZtomuDec and nbVXVYVZtomuRA (nbody/src/nbody_coordinates.c:226-287):
 * heliocentric position and velocity relative to the Sun rotated into the
 * equatorial frame, the velocity's components along the unit vectors of
 * increasing latitude and longitude over the distance, in milliarcseconds
 * per year (degrees * 3600 / 10^6 of the model's units). */
__device__ __forceinline__ void nbh_proper_motion(const PMFrame& f, double sun, double x,
                                                  double y, double z, double vx, double vy,
                                                  double vz, double& mudec, double& mura)
{
    const double px = x + sun, py = y, pz = z;
    const double ux = vx - f.sun_vel[0], uy = vy - f.sun_vel[1], uz = vz - f.sun_vel[2];
    const double X = f.rot[0][0] * px + f.rot[0][1] * py + f.rot[0][2] * pz;
    const double Y = f.rot[1][0] * px + f.rot[1][1] * py + f.rot[1][2] * pz;
    const double Z = f.rot[2][0] * px + f.rot[2][1] * py + f.rot[2][2] * pz;
    const double VX = f.rot[0][0] * ux + f.rot[0][1] * uy + f.rot[0][2] * uz;
    const double VY = f.rot[1][0] * ux + f.rot[1][1] * uy + f.rot[1][2] * uz;
    const double VZ = f.rot[2][0] * ux + f.rot[2][1] * uy + f.rot[2][2] * uz;
    double l = atan2(Y, X);
    const double b = atan2(Z, sqrt(X * X + Y * Y));
    if (l < 0.0) l += 2.0 * M_PI;
    /* findLatUnitVec goes through degrees: b' = 90 - b (degrees) */
    const double bdeg = b * 180.0 / M_PI;
    const double b2 = (90.0 - bdeg) * M_PI / 180.0;
    const double edx = -cos(l) * cos(b2), edy = -cos(b2) * sin(l), edz = sin(b2);
    const double erx = -sin(l), ery = cos(l);
    const double dist = sqrt(X * X + Y * Y + Z * Z);
    const double cdec = VX * edx + VY * edy + VZ * edz;
    const double cra = VX * erx + VY * ery;   /* e_l has no z component */
    mudec = (cdec / dist) * 180.0 / M_PI * 3600.0 / 1000000.0;
    mura = (cra / dist) * 180.0 / M_PI * 3600.0 / 1000000.0;
}

__global__ __launch_bounds__(256) void k_disp_values(int off, int cnt,
                                                     const double* __restrict__ px,
                                                     const double* __restrict__ py,
                                                     const double* __restrict__ pz,
                                                     const double* __restrict__ vx,
                                                     const double* __restrict__ vy,
                                                     const double* __restrict__ vz,
                                                     const unsigned char* __restrict__ ignore,
                                                     HistParams h, PMFrame f,
                                                     int* __restrict__ dbin,
                                                     double* __restrict__ dval)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= cnt) return;
    const int i = off + k;
    int bin = -1;
    double beta = 0.0, vl = 0.0, dist = 0.0, mudec = 0.0, mura = 0.0;
    if (!(ignore && ignore[i])) {
        const double x = px[i], y = py[i], z = pz[i];
        bin = nbh_lambda_beta_bin(h, x, y, z, beta);
        const double xs = x + h.sun;
        const double mag = sqrt(xs * xs + y * y + z * z);
        vl = (xs * vx[i] + y * vy[i] + z * vz[i]) / mag;
        dist = mag;
        nbh_proper_motion(f, h.sun, x, y, z, vx[i], vy[i], vz[i], mudec, mura);
    }
    dbin[k] = bin;
    dval[k] = beta;
    dval[(size_t) cnt + k] = vl;
    dval[2 * (size_t) cnt + k] = dist;
    dval[3 * (size_t) cnt + k] = mudec;
    dval[4 * (size_t) cnt + k] = mura;
}

__global__ __launch_bounds__(256) void k_disp_part(int cnt, int nbin, int initial, DispRound r,
                                                   const int* __restrict__ dbin,
                                                   const double* __restrict__ dval,
                                                   const double* __restrict__ state,
                                                   double* __restrict__ part)
{
    __shared__ int sb[NBH_DISP_TILE];
    __shared__ double sv[NBH_DISP_Q][NBH_DISP_TILE];
    const int base = blockIdx.x * NBH_DISP_TILE;
    const int m = min(NBH_DISP_TILE, cnt - base);
    for (int j = threadIdx.x; j < m; j += 256) {
        sb[j] = dbin[base + j];
        for (int q = 0; q < NBH_DISP_Q; ++q) sv[q][j] = dval[(size_t) q * cnt + base + j];
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbin; b += 256) {
        const double* st = state + (size_t) b * NBH_DISP_S;
        /* previous round: mean = sum / (count - removed), sigma = dispersion */
        double ave[NBH_DISP_Q], cut[NBH_DISP_Q];
        for (int q = 0; q < NBH_DISP_Q; ++q) {
            ave[q] = initial ? 0.0 : st[6 * q] / (st[NBH_DISP_RAW] - st[6 * q + 2]);
            cut[q] = initial ? 0.0 : r.sigma[q] * st[6 * q + 3];
        }
        double acc[NBH_DISP_W];
        for (int w = 0; w < NBH_DISP_W; ++w) acc[w] = 0.0;
        for (int j = 0; j < m; ++j) {
            if (sb[j] != b) continue;
            acc[NBH_DISP_W - 1] += 1.0;
            for (int q = 0; q < NBH_DISP_Q; ++q) {
                const double v = sv[q][j];
                if (initial || fabs(ave[q] - v) < cut[q]) {
                    acc[3 * q] += v;
                    acc[3 * q + 1] += v * v;
                } else {
                    acc[3 * q + 2] += 1.0;
                }
            }
        }
        double* o = part + ((size_t) blockIdx.x * nbin + b) * NBH_DISP_W;
        for (int w = 0; w < NBH_DISP_W; ++w) o[w] = acc[w];
    }
}

/* nbCalcDisp for one bin: sum, sq_sum, outliers removed, raw count -> the
 * dispersion and its error, kept from the previous round below 11 bodies. */
__device__ __forceinline__ void nbh_calc_disp(double sum, double sq, double removed,
                                              double raw, int initial, double correction,
                                              double& var, double& err)
{
    const double count = raw - removed;
    if (count > 10.0) {
        const double n_new = count - 1.0;
        const double n_ratio = count / n_new;
        const double mean = sum / count;
        double dispsq = (sq / n_new) - n_ratio * (mean * mean);
        if (!initial) dispsq *= correction;
        var = sqrt(dispsq);
        err = sqrt((count + 1) / (count * n_new)) * var;
    }
}

__global__ __launch_bounds__(256) void k_disp_finish(int nbin, int nblk, int initial, DispRound r,
                                                     const double* __restrict__ part,
                                                     double* __restrict__ state)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= nbin) return;
    double t[NBH_DISP_W];
    for (int w = 0; w < NBH_DISP_W; ++w) t[w] = 0.0;
    for (int k = 0; k < nblk; ++k) {
        const double* o = part + ((size_t) k * nbin + b) * NBH_DISP_W;
        for (int w = 0; w < NBH_DISP_W; ++w) t[w] += o[w];
    }
    double* st = state + (size_t) b * NBH_DISP_S;
    st[NBH_DISP_RAW] = t[NBH_DISP_W - 1];
    for (int q = 0; q < NBH_DISP_Q; ++q) {
        double* sq = st + 6 * q;
        /* every bin starts at dispersion 0 (nbody_histogram.c:846-854) */
        if (initial) sq[3] = sq[4] = 0.0;
        sq[0] = t[3 * q];
        sq[1] = t[3 * q + 1];
        sq[2] = t[3 * q + 2];
        nbh_calc_disp(sq[0], sq[1], sq[2], st[NBH_DISP_RAW], initial, r.correct[q], sq[3],
                      sq[4]);
    }
}

/* Histograms 1-7 from the final state: dispersions of beta and v_los, and
 * the averages of v_los, beta, distance, mu_dec and mu_ra with error
 * dispersion / sqrt(n) (nbody/src/nbody_histogram.c:996-1110; -1 and 0 below
 * 11 bodies).  out[(2 (k - 1) + {0: value, 1: error}) * nbin + b] for
 * histogram k. */
__global__ __launch_bounds__(256) void k_disp_out(int nbin, const double* __restrict__ state,
                                                  double* __restrict__ out)
{
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= nbin) return;
    const double* st = state + (size_t) b * NBH_DISP_S;
    const size_t N = (size_t) nbin;
    out[0 * N + b] = st[3];                         /* 1: beta dispersion */
    out[1 * N + b] = st[4];
    out[2 * N + b] = st[6 + 3];                     /* 2: v_los dispersion */
    out[3 * N + b] = st[6 + 4];
    /* 3: v_los, 4: beta, 5: distance, 6: mu_dec, 7: mu_ra */
    const int avg_q[5] = { 1, 0, 2, 3, 4 };
    for (int k = 0; k < 5; ++k) {
        const double* sq = st + 6 * avg_q[k];
        const int denom = (int) (st[NBH_DISP_RAW] - sq[2]);
        double v = 0.0, e = -1.0;
        if (denom > 10) {
            e = sq[3] / sqrt((double) denom);
            v = sq[0] / denom;
        }
        out[(4 + 2 * k) * N + b] = v;
        out[(5 + 2 * k) * N + b] = e;
    }
}

static inline int nbh_blocks(int n, int per) { return n > 0 ? (n + per - 1) / per : 1; }

static HistParams nbh_hist_params(const nbhip_histogram_params* hp);

void nbh_dispersion(nbhip_state* st, int off, int cnt, const nbhip_histogram_params* hp,
                    const nbhip_disp_params* dp, int* dbin, double* dval, double* part,
                    double* state, double* out)
{
    const HistParams h = nbh_hist_params(hp);
    const int nbin = hp->lambda_bins * hp->beta_bins;
    const int nblk = nbh_blocks(cnt, NBH_DISP_TILE);
    DispRound r;
    r.sigma[0] = dp->beta_sigma;
    r.sigma[1] = dp->vel_sigma;
    r.sigma[2] = dp->dist_sigma;
    r.sigma[3] = r.sigma[4] = dp->pm_sigma;
    r.correct[0] = dp->beta_correct;
    r.correct[1] = dp->vel_correct;
    r.correct[2] = dp->dist_correct;
    r.correct[3] = r.correct[4] = dp->pm_correct;
    /* cartesianalign's rotation (nbody/src/nbody_coordinates.c:192-224) */
    PMFrame f;
    const double cl = cos(M_PI - dp->l_ncp), sl = sin(M_PI - dp->l_ncp);
    const double cd = cos(M_PI_2 - dp->ngp_dec), sd = sin(M_PI_2 - dp->ngp_dec);
    const double ca = cos(dp->ngp_ra), sa = sin(dp->ngp_ra);
    f.rot[0][0] = ca * cd * cl - sa * sl; f.rot[0][1] = -ca * cd * sl - sa * cl; f.rot[0][2] = ca * sd;
    f.rot[1][0] = sa * cd * cl + ca * sl; f.rot[1][1] = -sa * cd * sl + ca * cl; f.rot[1][2] = sa * sd;
    f.rot[2][0] = -sd * cl; f.rot[2][1] = sd * sl; f.rot[2][2] = cd;
    for (int k = 0; k < 3; ++k) f.sun_vel[k] = dp->sun_vel[k];
    hipStream_t s = st->stream;
    hipLaunchKernelGGL(k_disp_values, dim3(nbh_blocks(cnt, 256)), dim3(256), 0, s, off, cnt,
                       st->px, st->py, st->pz, st->vx, st->vy, st->vz, st->ignore, h, f, dbin,
                       dval);
    for (int round = 0; round <= dp->iter_max; ++round) {
        const int initial = round == 0;
        hipLaunchKernelGGL(k_disp_part, dim3(nblk), dim3(256), 0, s, cnt, nbin, initial, r,
                           dbin, dval, state, part);
        hipLaunchKernelGGL(k_disp_finish, dim3(nbh_blocks(nbin, 256)), dim3(256), 0, s, nbin,
                           nblk, initial, r, part, state);
    }
    hipLaunchKernelGGL(k_disp_out, dim3(nbh_blocks(nbin, 256)), dim3(256), 0, s, nbin, state,
                       out);
}

static HistParams nbh_hist_params(const nbhip_histogram_params* hp)
{
    HistParams h;
    const double d2r = M_PI / 180.0;
    const double f = hp->phi * d2r, t = hp->theta * d2r, s = hp->psi * d2r;
    const double cf = cos(f), sf = sin(f), ct = cos(t), sn = sin(t), cs = cos(s), ss = sin(s);
    h.r[0][0] = cs * cf - ct * sf * ss; h.r[0][1] = cs * sf + ct * cf * ss; h.r[0][2] = ss * sn;
    h.r[1][0] = -ss * cf - ct * sf * cs; h.r[1][1] = -ss * sf + ct * cf * cs; h.r[1][2] = cs * sn;
    h.r[2][0] = sn * sf; h.r[2][1] = -sn * cf; h.r[2][2] = ct;
    h.sun = hp->sun_gc_dist;
    h.lstart = hp->lambda_start;
    h.lsize = (hp->lambda_end - hp->lambda_start) / (double) hp->lambda_bins;
    h.bstart = hp->beta_start;
    h.bsize = (hp->beta_end - hp->beta_start) / (double) hp->beta_bins;
    h.lbins = hp->lambda_bins;
    h.bbins = hp->beta_bins;
    return h;
}

void nbh_histogram(nbhip_state* st, const nbhip_histogram_params* hp, unsigned* d_counts,
                   unsigned* d_totals)
{
    const HistParams h = nbh_hist_params(hp);
    hipLaunchKernelGGL(k_histogram, dim3(nbh_blocks(st->n, 256)), dim3(256), 0, st->stream,
                       st->n, st->px, st->py, st->pz, st->ignore, st->simid, h, d_counts,
                       d_totals);
}

/* ---------------------------------------------------------------------- */
/* allocation-free histogram counts                                        */
/* ---------------------------------------------------------------------- */
/* Device scratch of nbhip_histogram_noalloc: the per-simulation bin counters
 * and totals.  One copy per GPU, held until the process ends; every call on
 * a GPU shares it, so calls that overlap on one GPU race.  In host code the
 * array's name stands for the symbol, not a device address: kernels receive
 * the per-device address that nbh_scratch_init looks up. */
static __device__ __align__(256) unsigned scratch[NBHIP_NOALLOC_MAX_COUNTERS];

#define NBH_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static unsigned* g_scratch_ptr[NBH_MAX_DEVICES];

/* Looks up this GPU's copy of the scratch array once; called by nbhip_create
 * and nbhip_batch_create with `device` current. */
unsigned nbh_scratch_init(int device)
{
    if (device < 0 || device >= NBH_MAX_DEVICES) {
        nbh_set_error("device %d: at most %d devices", device, NBH_MAX_DEVICES);
        return NBHIP_CAPABILITY_ERROR;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return NBHIP_SUCCESS;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        nbh_set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                      hipGetErrorString(e));
        return NBHIP_GPU_ERROR;
    }
    g_scratch_ptr[device] = static_cast<unsigned*>(p);
    return NBHIP_SUCCESS;
}

/* Fill kernel (not hipMemsetAsync, so the call also works under capture). */
__global__ __launch_bounds__(256) void k_fill_u32(unsigned* __restrict__ p, size_t n, unsigned v)
{
    for (size_t i = (size_t) blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t) gridDim.x * 256)
        p[i] = v;
}

extern "C" unsigned nbhip_histogram_noalloc(nbhip_state* st, const nbhip_histogram_params* hp,
                                            unsigned* counts, unsigned* totals)
{
    if (!st || !hp) return NBHIP_USER_ERROR;
    if (hp->lambda_bins < 1 || hp->beta_bins < 1 || !(hp->lambda_end > hp->lambda_start) ||
        !(hp->beta_end > hp->beta_start)) {
        nbh_set_error("nbhip_histogram_noalloc: empty bin range");
        return NBHIP_USER_ERROR;
    }
    const size_t nbin = (size_t) hp->lambda_bins * (size_t) hp->beta_bins;
    const size_t need = (size_t) st->nsim * (nbin + 1);
    if (nbin > NBHIP_NOALLOC_MAX_COUNTERS || need > NBHIP_NOALLOC_MAX_COUNTERS) {
        nbh_set_error("nbhip_histogram_noalloc: %d simulations x (%zu bins + 1) = %zu counters "
                      "exceed the scratch of %u",
                      st->nsim, nbin, need, (unsigned) NBHIP_NOALLOC_MAX_COUNTERS);
        return NBHIP_USER_ERROR;
    }
    NBH_CHECK(hipSetDevice(st->device));
    unsigned* base = (st->device >= 0 && st->device < NBH_MAX_DEVICES)
                         ? g_scratch_ptr[st->device] : nullptr;
    if (!base) {
        nbh_set_error("nbhip_histogram_noalloc: no scratch address for device %d", st->device);
        return NBHIP_GPU_ERROR;
    }
    unsigned* d_counts = base;
    unsigned* d_totals = base + (size_t) st->nsim * nbin;
    /* the previous call's counts are still there: clear what this one uses */
    const int gfill = (int) std::min<size_t>((need + 255) / 256, 1024);
    hipLaunchKernelGGL(k_fill_u32, dim3(gfill), dim3(256), 0, st->stream, base, need, 0u);
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
