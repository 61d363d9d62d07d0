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
/* beta and line-of-sight velocity dispersions per bin                     */
/* ---------------------------------------------------------------------- */
/* nbCreateHistogram's dispersion histograms 1 (beta) and 2 (line-of-sight
 * velocity) with the outlier rejection that follows
 * (nbody/src/nbody_histogram.c:858-1000; nbCalcDisp and nbRemoveOutliers,
 * nbody/src/nbody_mass.c:195-316; calc_vLOS :171-179), for the bodies of
 * one simulation.
 *
 * k_disp_values records each body's bin (-1: ignored or outside), beta and
 * line-of-sight velocity once.  Each round of sums is k_disp_part, one
 * workgroup per 1024 consecutive bodies with a thread per bin that adds the
 * workgroup's bodies in body order, then k_disp_finish, a thread per bin
 * that adds the workgroups' partial sums in workgroup order and applies
 * nbCalcDisp.  The sums are therefore deterministic (independent of timing
 * and batching) and differ from the reference's single running sum only in
 * association.  Round 0 takes every binned body; each later round keeps a
 * body when |mean - value| < sigma * dispersion of the previous round. */
#define NBH_DISP_TILE 1024
#define NBH_DISP_W 7          /* per bin: beta sum, sq, removed; vlos sum, sq, removed; count */
#define NBH_DISP_S 12         /* per bin state: for beta and vlos: sum sq removed var err pad */

__global__ __launch_bounds__(256) void k_disp_values(int off, int cnt,
                                                     const double* __restrict__ px,
                                                     const double* __restrict__ py,
                                                     const double* __restrict__ pz,
                                                     const double* __restrict__ vx,
                                                     const double* __restrict__ vy,
                                                     const double* __restrict__ vz,
                                                     const unsigned char* __restrict__ ignore,
                                                     HistParams h, int* __restrict__ dbin,
                                                     double* __restrict__ dbeta,
                                                     double* __restrict__ dvlos)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= cnt) return;
    const int i = off + k;
    int bin = -1;
    double beta = 0.0, vl = 0.0;
    if (!(ignore && ignore[i])) {
        const double x = px[i], y = py[i], z = pz[i];
        bin = nbh_lambda_beta_bin(h, x, y, z, beta);
        const double xs = x + h.sun;
        const double mag = sqrt(xs * xs + y * y + z * z);
        vl = (xs * vx[i] + y * vy[i] + z * vz[i]) / mag;
    }
    dbin[k] = bin;
    dbeta[k] = beta;
    dvlos[k] = vl;
}

__global__ __launch_bounds__(256) void k_disp_part(int cnt, int nbin, int initial,
                                                   double beta_sigma, double vel_sigma,
                                                   const int* __restrict__ dbin,
                                                   const double* __restrict__ dbeta,
                                                   const double* __restrict__ dvlos,
                                                   const double* __restrict__ state,
                                                   double* __restrict__ part)
{
    __shared__ int sb[NBH_DISP_TILE];
    __shared__ double sbeta[NBH_DISP_TILE];
    __shared__ double svl[NBH_DISP_TILE];
    const int base = blockIdx.x * NBH_DISP_TILE;
    const int m = min(NBH_DISP_TILE, cnt - base);
    for (int j = threadIdx.x; j < m; j += 256) {
        sb[j] = dbin[base + j];
        sbeta[j] = dbeta[base + j];
        svl[j] = dvlos[base + j];
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nbin; b += 256) {
        const double* st = state + (size_t) b * NBH_DISP_S;
        /* previous round: mean = sum / (count - removed), sigma = dispersion */
        double cB = 0.0, aveB = 0.0, sigB = 0.0, aveV = 0.0, sigV = 0.0;
        if (!initial) {
            cB = st[11];
            aveB = st[0] / (cB - st[2]);
            sigB = st[3];
            aveV = st[5] / (cB - st[7]);
            sigV = st[8];
        }
        double s0 = 0.0, q0 = 0.0, r0 = 0.0, s1 = 0.0, q1 = 0.0, r1 = 0.0, c = 0.0;
        for (int j = 0; j < m; ++j) {
            if (sb[j] != b) continue;
            c += 1.0;
            const double vb = sbeta[j], vv = svl[j];
            if (initial || fabs(aveB - vb) < beta_sigma * sigB) {
                s0 += vb;
                q0 += vb * vb;
            } else {
                r0 += 1.0;
            }
            if (initial || fabs(aveV - vv) < vel_sigma * sigV) {
                s1 += vv;
                q1 += vv * vv;
            } else {
                r1 += 1.0;
            }
        }
        double* o = part + ((size_t) blockIdx.x * nbin + b) * NBH_DISP_W;
        o[0] = s0; o[1] = q0; o[2] = r0; o[3] = s1; o[4] = q1; o[5] = r1; o[6] = c;
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

__global__ __launch_bounds__(256) void k_disp_finish(int nbin, int nblk, int initial,
                                                     double beta_correct, double vel_correct,
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
    if (initial) {
        /* every bin starts at dispersion 0 (nbody_histogram.c:846-854) */
        st[3] = st[4] = st[8] = st[9] = 0.0;
    }
    st[0] = t[0]; st[1] = t[1]; st[2] = t[2];
    st[5] = t[3]; st[6] = t[4]; st[7] = t[5];
    st[11] = t[6];
    nbh_calc_disp(t[0], t[1], t[2], t[6], initial, beta_correct, st[3], st[4]);
    nbh_calc_disp(t[3], t[4], t[5], t[6], initial, vel_correct, st[8], st[9]);
}

static inline int nbh_blocks(int n, int per) { return n > 0 ? (n + per - 1) / per : 1; }

static HistParams nbh_hist_params(const nbhip_histogram_params* hp);

void nbh_dispersion(nbhip_state* st, int off, int cnt, const nbhip_histogram_params* hp,
                    const nbhip_disp_params* dp, int* dbin, double* dbeta, double* dvlos,
                    double* part, double* state)
{
    const HistParams h = nbh_hist_params(hp);
    const int nbin = hp->lambda_bins * hp->beta_bins;
    const int nblk = nbh_blocks(cnt, NBH_DISP_TILE);
    hipStream_t s = st->stream;
    hipLaunchKernelGGL(k_disp_values, dim3(nbh_blocks(cnt, 256)), dim3(256), 0, s, off, cnt,
                       st->px, st->py, st->pz, st->vx, st->vy, st->vz, st->ignore, h, dbin,
                       dbeta, dvlos);
    for (int round = 0; round <= dp->iter_max; ++round) {
        const int initial = round == 0;
        hipLaunchKernelGGL(k_disp_part, dim3(nblk), dim3(256), 0, s, cnt, nbin, initial,
                           dp->beta_sigma, dp->vel_sigma, dbin, dbeta, dvlos, state, part);
        hipLaunchKernelGGL(k_disp_finish, dim3(nbh_blocks(nbin, 256)), dim3(256), 0, s, nbin,
                           nblk, initial, dp->beta_correct, dp->vel_correct, part, state);
    }
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
