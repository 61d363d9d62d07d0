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
        const double X = px[i] + h.sun, Y = py[i], Z = pz[i];
        const double tx = h.r[0][0] * X + h.r[0][1] * Y + h.r[0][2] * Z;
        const double ty = h.r[1][0] * X + h.r[1][1] * Y + h.r[1][2] * Z;
        const double tz = -(h.r[2][0] * X + h.r[2][1] * Y + h.r[2][2] * Z);
        const double r = sqrt(tx * tx + ty * ty + tz * tz);
        const double lambda = atan2(ty, tx) * (180.0 / M_PI);
        const double beta = asin(tz / r) * (180.0 / M_PI);
        const double li = floor((lambda - h.lstart) / h.lsize);
        const double bi = floor((beta - h.bstart) / h.bsize);
        if (li >= 0.0 && li < (double) h.lbins && bi >= 0.0 && bi < (double) h.bbins)
            bin = (int) li * h.bbins + (int) bi;
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

static inline int nbh_blocks(int n, int per) { return n > 0 ? (n + per - 1) / per : 1; }

void nbh_histogram(nbhip_state* st, const nbhip_histogram_params* hp, unsigned* d_counts,
                   unsigned* d_totals)
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
    hipLaunchKernelGGL(k_histogram, dim3(nbh_blocks(st->n, 256)), dim3(256), 0, st->stream,
                       st->n, st->px, st->py, st->pz, st->ignore, st->simid, h, d_counts,
                       d_totals);
}
