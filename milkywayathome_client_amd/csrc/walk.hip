/*
 * Force walk on gfx950.
 *
 * Each wavefront owns 64 consecutive bodies in Morton order and streams the
 * pre-order node table once, in order.  A lane keeps its own position lp in
 * its reference walk (nbGravity, nbody/src/nbody_grav.c:42-124): a lane whose
 * lp equals the wave's current node q tests it and either interacts
 * (lp = Next(q)) or opens it (lp = q + 1 = More(q)).  Lanes parked on a later
 * node sit out.  The wave moves to q + 1 if any lane opened q, otherwise to
 * Next(q); every parked lane waits at Next() of an ancestor of q, never
 * inside (q, Next(q)), so the wave's current node is always the minimum lp
 * and every lane sees exactly its reference interaction list in the
 * reference order.  Node records are wave-uniform, so they arrive through
 * scalar loads; only interactions spend vector ALU.
 */
#include "nbhip_device.h"

struct WalkArgs {
    int n;
    int kick;
    int use_pot;
    int pad;
    double eps2;
    double dtHalf;
    const NodeW* node;
    const double* nquad;
    const double4* spos;
    const unsigned* perm;
    double* oax;
    double* oay;
    double* oaz;
    double* vx;
    double* vy;
    double* vz;
    DevScalars* scal;
    nbhip_potential pot;
};

template <bool QUAD, bool COUNT>
__global__ __launch_bounds__(256) void k_walk64(WalkArgs a)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (a.scal->status & (NBHIP_CELL_OVERFLOW_ERROR | NBHIP_ERROR)) return;
    const int M = a.scal->nmassive;
    const int nn = a.scal->nnode;
    const bool valid = s < a.n;
    if (__ballot(valid) == 0ull) return;
    const double4 p = valid ? a.spos[s] : make_double4(0.0, 0.0, 0.0, 0.0);
    const int self = (valid && s < M) ? s : -2;
    int lp = valid ? 0 : nn;
    double ax = 0.0, ay = 0.0, az = 0.0;
    bool met = false;
    unsigned nbb = 0, nbc = 0, nop = 0;
    const double eps2 = a.eps2;

    int q = 0;
    while (q < nn) {
        q = __builtin_amdgcn_readfirstlane(q);
        const NodeW nd = a.node[q];
        bool open = false;
        if (lp == q) {
            const double dx = nd.x - p.x;
            const double dy = nd.y - p.y;
            const double dz = nd.z - p.z;
            double drSq = dz * dz + (dy * dy + dx * dx);
            if (nd.body >= 0 || drSq >= nd.rcrit2) {
                if (nd.body != self) {
                    drSq += eps2;
                    const double drab = sqrt(drSq);
                    const double phii = nd.m / drab;
                    const double mor3 = phii / drSq;
                    ax += mor3 * dx;
                    ay += mor3 * dy;
                    az += mor3 * dz;
                    if (QUAD && nd.body < 0) {
                        const double* Q = a.nquad + 6 * q;
                        const double qxx = Q[0], qxy = Q[1], qxz = Q[2];
                        const double qyy = Q[3], qyz = Q[4], qzz = Q[5];
                        const double qx = qxx * dx + qxy * dy + qxz * dz;
                        const double qy = qxy * dx + qyy * dy + qyz * dz;
                        const double qz = qxz * dx + qyz * dy + qzz * dz;
                        const double drQdr = qx * dx + qy * dy + qz * dz;
                        const double dr5inv = 1.0 / ((drSq * drSq) * drab);
                        const double phiQ = 2.5 * (dr5inv * drQdr) / drSq;
                        ax += phiQ * dx;
                        ay += phiQ * dy;
                        az += phiQ * dz;
                        ax -= dr5inv * qx;
                        ay -= dr5inv * qy;
                        az -= dr5inv * qz;
                    }
                    if (COUNT) {
                        if (nd.body >= 0) ++nbb; else ++nbc;
                    }
                } else {
                    met = true;
                }
                lp = nd.next;
            } else {
                lp = q + 1;
                open = true;
                if (COUNT) ++nop;
            }
        }
        const int qn = __ballot(open) ? q + 1 : nd.next;
        if (qn <= q) {
            /* a Next() link that does not advance: corrupt table, stop */
            if (__lane_id() == 0) atomicOr(&a.scal->status, NBHIP_ERROR);
            break;
        }
        q = qn;
    }

    if (COUNT) {
        unsigned long long c0 = nbh_wave_sum_u64(nbb);
        unsigned long long c1 = nbh_wave_sum_u64(nbc);
        unsigned long long c2 = nbh_wave_sum_u64(nop);
        if (__lane_id() == 0) {
            atomicAdd(&a.scal->counters[0], c0);
            atomicAdd(&a.scal->counters[1], c1);
            atomicAdd(&a.scal->counters[2], c2);
        }
    }
    if (!valid) return;
    if (!met) atomicOr(&a.scal->incest, 1u);
    if (a.use_pot) {
        double ex, ey, ez;
        nbh_ext_accel(a.pot, p.x, p.y, p.z, ex, ey, ez);
        ax = ax + ex;
        ay = ay + ey;
        az = az + ez;
    }
    const unsigned j = a.perm[s];
    a.oax[j] = ax;
    a.oay[j] = ay;
    a.oaz[j] = az;
    if (a.kick) {
        /* advanceVelocities (nbody/src/nbody_plain.c:301-315) */
        a.vx[j] = a.vx[j] + a.dtHalf * (ax + 0.0);
        a.vy[j] = a.vy[j] + a.dtHalf * (ay + 0.0);
        a.vz[j] = a.vz[j] + a.dtHalf * (az + 0.0);
    }
}

/* nbMapForceBody_Exact (nbody/src/nbody_grav.c:229-293): O(N^2) in body
 * order, one body per lane, sources broadcast across the wave. */
__global__ __launch_bounds__(256) void k_exact_all(int n, const double* __restrict__ px,
                                                   const double* __restrict__ py,
                                                   const double* __restrict__ pz,
                                                   const double* __restrict__ m, double eps2,
                                                   int use_pot, nbhip_potential pot, int kick,
                                                   double dtHalf, double* __restrict__ oax,
                                                   double* __restrict__ oay,
                                                   double* __restrict__ oaz,
                                                   double* __restrict__ vx,
                                                   double* __restrict__ vy,
                                                   double* __restrict__ vz)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < n;
    const double x0 = valid ? px[i] : 0.0, y0 = valid ? py[i] : 0.0, z0 = valid ? pz[i] : 0.0;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int k = 0; k < n; ++k) {
        const double dx = px[k] - x0;
        const double dy = py[k] - y0;
        const double dz = pz[k] - z0;
        const double drSq = dz * dz + (dy * dy + dx * dx) + eps2;
        const double drab = sqrt(drSq);
        const double phii = m[k] / drab;
        const double mor3 = phii / drSq;
        ax += mor3 * dx;
        ay += mor3 * dy;
        az += mor3 * dz;
    }
    if (!valid) return;
    if (use_pot) {
        double ex, ey, ez;
        nbh_ext_accel(pot, x0, y0, z0, ex, ey, ez);
        ax = ax + ex;
        ay = ay + ey;
        az = az + ez;
    }
    oax[i] = ax;
    oay[i] = ay;
    oaz[i] = az;
    if (kick) {
        vx[i] = vx[i] + dtHalf * (ax + 0.0);
        vy[i] = vy[i] + dtHalf * (ay + 0.0);
        vz[i] = vz[i] + dtHalf * (az + 0.0);
    }
}

/* FP64 direct-sum checker on sampled bodies: one block per sample, the
 * sum over sources split across 256 lanes and reduced in a fixed order. */
__global__ __launch_bounds__(256) void k_direct_sample(int n, const double* __restrict__ px,
                                                       const double* __restrict__ py,
                                                       const double* __restrict__ pz,
                                                       const double* __restrict__ m, double eps2,
                                                       const int* __restrict__ idx,
                                                       double* __restrict__ acc)
{
    const int b = idx[blockIdx.x];
    const double x0 = px[b], y0 = py[b], z0 = pz[b];
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) {
        const double dx = px[k] - x0;
        const double dy = py[k] - y0;
        const double dz = pz[k] - z0;
        const double drSq = dz * dz + (dy * dy + dx * dx) + eps2;
        const double drab = sqrt(drSq);
        const double mor3 = (m[k] / drab) / drSq;
        ax += mor3 * dx;
        ay += mor3 * dy;
        az += mor3 * dz;
    }
    __shared__ double red[3][256];
    red[0][threadIdx.x] = ax;
    red[1][threadIdx.x] = ay;
    red[2][threadIdx.x] = az;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int) threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
            red[2][threadIdx.x] += red[2][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        acc[3 * blockIdx.x + 0] = red[0][0];
        acc[3 * blockIdx.x + 1] = red[1][0];
        acc[3 * blockIdx.x + 2] = red[2][0];
    }
}

static inline int nbh_blocks(int n, int per) { return n > 0 ? (n + per - 1) / per : 1; }

void nbh_walk(nbhip_state* st, int kick, int count, double* oax, double* oay, double* oaz,
              int self_only)
{
    const nbhip_ctx& c = st->ctx;
    if (c.criterion == NBHIP_EXACT) {
        hipLaunchKernelGGL(k_exact_all, dim3(nbh_blocks(st->n, 256)), dim3(256), 0, st->stream,
                           st->n, st->px, st->py, st->pz, st->m, c.eps2,
                           self_only ? 0 : c.use_potential, c.pot, kick, 0.5 * c.timestep, oax,
                           oay, oaz, st->vx, st->vy, st->vz);
        return;
    }
    WalkArgs a;
    a.n = st->n;
    a.kick = kick;
    a.use_pot = self_only ? 0 : c.use_potential;
    a.pad = 0;
    a.eps2 = c.eps2;
    a.dtHalf = 0.5 * c.timestep;
    a.node = st->node;
    a.nquad = st->nquad;
    a.spos = st->spos;
    a.perm = st->perm;
    a.oax = oax;
    a.oay = oay;
    a.oaz = oaz;
    a.vx = st->vx;
    a.vy = st->vy;
    a.vz = st->vz;
    a.scal = st->scal;
    a.pot = c.pot;
    dim3 grid(nbh_blocks(st->n, 256)), block(256);
    if (c.use_quad) {
        if (count) hipLaunchKernelGGL((k_walk64<true, true>), grid, block, 0, st->stream, a);
        else hipLaunchKernelGGL((k_walk64<true, false>), grid, block, 0, st->stream, a);
    } else {
        if (count) hipLaunchKernelGGL((k_walk64<false, true>), grid, block, 0, st->stream, a);
        else hipLaunchKernelGGL((k_walk64<false, false>), grid, block, 0, st->stream, a);
    }
}

void nbh_direct(nbhip_state* st, int nsample, const int* d_idx, double* d_acc)
{
    if (nsample <= 0) return;
    hipLaunchKernelGGL(k_direct_sample, dim3(nsample), dim3(256), 0, st->stream, st->n, st->px,
                       st->py, st->pz, st->m, st->ctx.eps2, d_idx, d_acc);
}
