/*
 * Exact criterion (nbGravity_Exact / nbMapForceBody_Exact,
 * nbody/src/nbody_grav.c:205-293): the O(N^2) pairwise sum, forward and
 * adjoint.
 *
 * One lane per body; the sources of a simulation of n bodies are split into
 * nbh_exact_segments(n) consecutive segments, segment s summed by grid row
 * blockIdx.y = s.  A few thousand bodies then still fill the chip (20 000
 * bodies are only 313 wavefronts, one serial sum each, otherwise), and the
 * segmentation is a property of the body's own simulation, so batched and
 * single runs add in the same order.  A segment is summed in source order,
 * as the reference does, and k_exact_finish adds the segments in order; with
 * one segment (n < 4096) the result is the reference's sum bit for bit.
 *
 * A workgroup whose 256 bodies belong to one simulation stages the sources
 * through LDS, 256 at a time (read once per workgroup instead of once per
 * lane); a workgroup that straddles two simulations reads them per lane.
 * Both paths evaluate the same expressions in the same order.
 */
#include "nbhip_device.h"

#define NBH_EXACT_TILE 256

/* Source range of segment `seg` of simulation t (empty when seg is past the
 * simulation's segment count). */
__device__ __forceinline__ void nbh_exact_range(const SimTab& t, int seg, int& a, int& b)
{
    const int S = nbh_exact_segments(t.body_n);
    if (seg >= S) {
        a = b = 0;
        return;
    }
    a = t.body_off + (int) (((long long) seg * t.body_n) / S);
    b = t.body_off + (int) (((long long) (seg + 1) * t.body_n) / S);
}

/* Forward pair term: nbGravity_Exact's expression order. */
struct ExactFwd {
    double x0, y0, z0, eps2;
    double ax = 0.0, ay = 0.0, az = 0.0;
    __device__ __forceinline__ void add(double xk, double yk, double zk, double mk)
    {
        const double dx = xk - x0;
        const double dy = yk - y0;
        const double dz = zk - z0;
        const double drSq = dz * dz + (dy * dy + dx * dx) + eps2;
        const double drab = sqrt(drSq);
        const double phii = mk / drab;
        const double mor3 = phii / drSq;
        ax += mor3 * dx;
        ay += mor3 * dy;
        az += mor3 * dz;
    }
};

/* Adjoint pair term: (J^T g)_i = sum_k T(d_ik) (m_i g_k - m_k g_i),
 * dL/dm_i = -sum_k g_k . d_ik / s^3, d_ik = x_k - x_i,
 * T(d) = I / s^3 - 3 d d^T / s^5 (the self term vanishes identically). */
struct ExactAdj {
    double x0, y0, z0, eps2, mi, gi0, gi1, gi2;
    double ox = 0.0, oy = 0.0, oz = 0.0, om = 0.0;
    __device__ __forceinline__ void add(double xk, double yk, double zk, double mk, double gk0,
                                        double gk1, double gk2)
    {
        const double dx = xk - x0, dy = yk - y0, dz = zk - z0;
        const double s2 = dx * dx + dy * dy + dz * dz + eps2;
        const double s3 = 1.0 / (s2 * sqrt(s2));
        const double s5 = s3 / s2;
        const double w0 = mi * gk0 - mk * gi0;
        const double w1 = mi * gk1 - mk * gi1;
        const double w2 = mi * gk2 - mk * gi2;
        const double dw = dx * w0 + dy * w1 + dz * w2;
        ox += s3 * w0 - 3.0 * s5 * dw * dx;
        oy += s3 * w1 - 3.0 * s5 * dw * dy;
        oz += s3 * w2 - 3.0 * s5 * dw * dz;
        om -= (gk0 * dx + gk1 * dy + gk2 * dz) * s3;
    }
};

/* Partial sums of one segment: part[(seg * W + c) * n + i], W = 3 (forward
 * acceleration) or 4 (adjoint: position gradient, mass gradient).  ga is the
 * acceleration adjoint (3 arrays of n), adjoint only. */
template <bool ADJ>
__global__ __launch_bounds__(256) void k_exact_part(int n, const double* __restrict__ px,
                                                    const double* __restrict__ py,
                                                    const double* __restrict__ pz,
                                                    const double* __restrict__ m,
                                                    const int* __restrict__ simid,
                                                    const SimTab* __restrict__ sims,
                                                    const double* __restrict__ ga,
                                                    double* __restrict__ part)
{
    __shared__ double4 tile[NBH_EXACT_TILE];
    __shared__ double4 tileg[ADJ ? NBH_EXACT_TILE : 1];
    __shared__ int s_mixed;
    const int tid = threadIdx.x;
    const int i = blockIdx.x * 256 + tid;
    const int seg = blockIdx.y;
    const bool valid = i < n;
    const size_t N = (size_t) n;
    const int sim = valid ? simid[i] : -1;
    const int sim0 = simid[blockIdx.x * 256];
    if (tid == 0) s_mixed = 0;
    __syncthreads();
    if (valid && sim != sim0) s_mixed = 1;       /* benign race: every writer stores 1 */
    __syncthreads();
    const bool uniform = s_mixed == 0;

    int a = 0, b = 0;
    double eps2 = 0.0;
    if (valid) {
        const SimTab& t = sims[sim];
        nbh_exact_range(t, seg, a, b);
        eps2 = t.eps2;
    }
    const double x0 = valid ? px[i] : 0.0, y0 = valid ? py[i] : 0.0, z0 = valid ? pz[i] : 0.0;
    ExactFwd F;
    ExactAdj G;
    if (ADJ) {
        G.x0 = x0; G.y0 = y0; G.z0 = z0; G.eps2 = eps2;
        G.mi = valid ? m[i] : 0.0;
        G.gi0 = valid ? ga[i] : 0.0;
        G.gi1 = valid ? ga[N + i] : 0.0;
        G.gi2 = valid ? ga[2 * N + i] : 0.0;
    } else {
        F.x0 = x0; F.y0 = y0; F.z0 = z0; F.eps2 = eps2;
    }

    if (uniform) {
        /* every body of the workgroup sums the same segment */
        int ua, ub;
        nbh_exact_range(sims[sim0], seg, ua, ub);
        if (ua >= ub) return;                     /* whole workgroup: no barrier follows */
        for (int base = ua; base < ub; base += NBH_EXACT_TILE) {
            const int k = base + tid;
            if (k < ub) {
                tile[tid] = make_double4(px[k], py[k], pz[k], m[k]);
                if (ADJ) tileg[tid] = make_double4(ga[k], ga[N + k], ga[2 * N + k], 0.0);
            }
            __syncthreads();
            const int cnt = min(NBH_EXACT_TILE, ub - base);
            for (int j = 0; j < cnt; ++j) {
                const double4 s = tile[j];
                if (ADJ) {
                    const double4 gk = tileg[j];
                    G.add(s.x, s.y, s.z, s.w, gk.x, gk.y, gk.z);
                } else {
                    F.add(s.x, s.y, s.z, s.w);
                }
            }
            __syncthreads();
        }
    } else {
        for (int k = a; k < b; ++k) {
            if (ADJ)
                G.add(px[k], py[k], pz[k], m[k], ga[k], ga[N + k], ga[2 * N + k]);
            else
                F.add(px[k], py[k], pz[k], m[k]);
        }
    }
    if (!valid || a >= b) return;
    constexpr int W = ADJ ? 4 : 3;
    double* p = part + (size_t) seg * W * N + i;
    if (ADJ) {
        p[0] = G.ox;
        p[N] = G.oy;
        p[2 * N] = G.oz;
        p[3 * N] = G.om;
    } else {
        p[0] = F.ax;
        p[N] = F.ay;
        p[2 * N] = F.az;
    }
}

/* Segments added in order, then the external potential and the optional
 * closing half-kick (forward; nbMapForceBody_Exact, nbody/src/nbody_grav.c:
 * 229-293) or the potential's Jacobian term and the accumulation into the
 * gradient (adjoint). */
template <bool ADJ>
__global__ __launch_bounds__(256) void k_exact_finish(int n, const double* __restrict__ px,
                                                      const double* __restrict__ py,
                                                      const double* __restrict__ pz,
                                                      const int* __restrict__ simid,
                                                      const SimTab* __restrict__ sims,
                                                      const double* __restrict__ part,
                                                      int use_pot, nbhip_potential pot,
                                                      int kick, double dtHalf,
                                                      double* __restrict__ oax,
                                                      double* __restrict__ oay,
                                                      double* __restrict__ oaz,
                                                      double* __restrict__ vx,
                                                      double* __restrict__ vy,
                                                      double* __restrict__ vz,
                                                      double* __restrict__ g,
                                                      const DevScalars* __restrict__ scal)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n || scal->halt) return;
    const size_t N = (size_t) n;
    constexpr int W = ADJ ? 4 : 3;
    const int S = nbh_exact_segments(sims[simid[i]].body_n);
    double s[W];
    for (int c = 0; c < W; ++c) s[c] = part[c * N + i];
    for (int seg = 1; seg < S; ++seg)
        for (int c = 0; c < W; ++c) s[c] += part[((size_t) seg * W + c) * N + i];
    const double x0 = px[i], y0 = py[i], z0 = pz[i];
    if (ADJ) {
        const double* ga = g + 6 * N;
        if (use_pot) {
            double ex, ey, ez;
            nbh_ext_jac(pot, x0, y0, z0, ga[i], ga[N + i], ga[2 * N + i], ex, ey, ez);
            s[0] += ex;
            s[1] += ey;
            s[2] += ez;
        }
        g[i] += s[0];
        g[N + i] += s[1];
        g[2 * N + i] += s[2];
        g[9 * N + i] += s[3];
        return;
    }
    double ax = s[0], ay = s[1], az = s[W - 1];
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

static inline int nbh_blocks(int n, int per) { return n > 0 ? (n + per - 1) / per : 1; }

void nbh_exact_force(nbhip_state* st, int kick, double* oax, double* oay, double* oaz,
                     int use_pot)
{
    if (st->n <= 0) return;
    const nbhip_ctx& c = st->ctx;
    hipStream_t s = st->stream;
    const int nb = nbh_blocks(st->n, 256);
    hipLaunchKernelGGL(k_exact_part<false>, dim3(nb, st->exact_seg), dim3(256), 0, s, st->n,
                       st->px, st->py, st->pz, st->m, st->simid, st->sims, nullptr,
                       st->exact_part);
    hipLaunchKernelGGL(k_exact_finish<false>, dim3(nb), dim3(256), 0, s, st->n, st->px, st->py,
                       st->pz, st->simid, st->sims, st->exact_part, use_pot, c.pot, kick,
                       0.5 * c.timestep, oax, oay, oaz, st->vx, st->vy, st->vz, nullptr,
                       st->scal);
}

void nbh_exact_adjoint(nbhip_state* st, double* g)
{
    if (st->n <= 0) return;
    const nbhip_ctx& c = st->ctx;
    hipStream_t s = st->stream;
    const int nb = nbh_blocks(st->n, 256);
    const size_t N = (size_t) st->n;
    hipLaunchKernelGGL(k_exact_part<true>, dim3(nb, st->exact_seg), dim3(256), 0, s, st->n,
                       st->px, st->py, st->pz, st->m, st->simid, st->sims, g + 6 * N,
                       st->exact_part);
    hipLaunchKernelGGL(k_exact_finish<true>, dim3(nb), dim3(256), 0, s, st->n, st->px, st->py,
                       st->pz, st->simid, st->sims, st->exact_part, c.use_potential, c.pot, 0,
                       0.0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, g, st->scal);
}
