/*
 * Adjoint of the KDK leapfrog (advancePosVel / advanceVelocities,
 * nbody/src/nbody_plain.c:268-315, around nbGravity, nbody/src/nbody_grav.c)
 * with the tree of each step held fixed: the cells, their membership and
 * every opening decision are those of the forward step, and the derivative
 * is taken of the forces as the walk evaluates them -- monopole and
 * quadrupole terms of accepted cells, whose centres of mass, masses and
 * quadrupoles are themselves functions of the member bodies.
 *
 * One force adjoint, given g = dL/da for every body:
 *   1. k_adj_walk replays each body's walk (same lanes, same node stream,
 *      same accept/open tests as the forward FP64 walk).  For an accepted
 *      node with d = cm - x, s^2 = |d|^2 + eps2:
 *        a    = M d / s^3 + 2.5 (d.Q d) d / s^7 - Q d / s^5
 *        H g  = (da/dd) g      (symmetric 3x3, so also (da/dd)^T g)
 *      the body's own position collects -H g; the node collects +H g
 *      towards its centre of mass, (d.g)/s^3 towards its mass and the
 *      derivative with respect to its six quadrupole components.  Lanes
 *      that used the node are summed across the wave and one lane adds
 *      the sums to the node's adjoint record.
 *   2. k_adj_down walks the cell levels from the roots down and hands each
 *      cell's adjoint to its children through the summarisation formulas
 *        cm = sum m_c cm_c / M,  M = sum m_c,
 *        Q  = sum m_c (3 d_c d_c^T - |d_c|^2 I) + Q_c,  d_c = cm_c - cm
 *      (the dependence of Q on cm drops out because sum m_c d_c = 0).
 *   3. k_adj_body adds each leaf's record to its body and, inside a step,
 *      applies the adjoint of the drift and the opening half-kick.
 * The external potential adds J_ext g per body (nbh_ext_jac).  The Exact
 * criterion uses the O(N^2) form (exact.hip).
 */
#include "nbhip_device.h"

struct AdjWalk {
    int n;
    int use_pot;
    int multi;
    int cap;
    nbhip_potential pot;
};

/* dL/d(acceleration) of one accepted node for one lane.  Returns the vector
 * H g and the mass / quadrupole parts of the node's adjoint. */
template <bool QUAD>
__global__ __launch_bounds__(256) void k_adj_walk(const NodeW* __restrict__ node,
                                                  const double* __restrict__ nquad,
                                                  const double4* __restrict__ spos,
                                                  const unsigned* __restrict__ perm,
                                                  const int* __restrict__ simid_s,
                                                  const SimTab* __restrict__ sims,
                                                  const double* __restrict__ gax,
                                                  const double* __restrict__ gay,
                                                  const double* __restrict__ gaz,
                                                  double* __restrict__ gxx,
                                                  double* __restrict__ gxy,
                                                  double* __restrict__ gxz,
                                                  double* __restrict__ gnode,
                                                  const DevScalars* __restrict__ scal, AdjWalk o)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    if ((scal->status & (NBHIP_CELL_OVERFLOW_ERROR | NBHIP_ERROR)) || scal->halt) return;
    const int M = scal->nmassive;
    const bool valid = s < o.n;
    if (__ballot(valid) == 0ull) return;
    const double4 p = valid ? spos[s] : make_double4(0.0, 0.0, 0.0, 0.0);
    const unsigned j = valid ? perm[s] : 0u;
    const double g0 = valid ? gax[j] : 0.0;
    const double g1 = valid ? gay[j] : 0.0;
    const double g2 = valid ? gaz[j] : 0.0;
    const LaneStart L = nbh_lane_start(s, valid, M, simid_s, sims);
    const int self = L.self, my_end = L.end;
    int lp = L.lp;
    const double eps2 = L.eps2;
    const int cap = o.cap;
    double ox = 0.0, oy = 0.0, oz = 0.0;

    int q = L.q0;
    while (q < L.qend) {
        const NodeW nd = node[q];
        const int nd_next = __builtin_amdgcn_readfirstlane(nd.next);
        const bool cell = __builtin_amdgcn_readfirstlane(nd.body) < 0;
        const bool active = lp == q;
        const double dx = nd.x - p.x;
        const double dy = nd.y - p.y;
        const double dz = nd.z - p.z;
        const double drSq = dz * dz + (dy * dy + dx * dx);
        const bool accept = active && (!cell || drSq >= nd.rcrit2);
        const bool open = active && !accept;
        const bool any_open = __builtin_amdgcn_ballot_w64(open) != 0ull;
        const int qn = any_open ? q + 1 : nd_next;
        if (qn <= q) break;          /* corrupt table; the forward walk reports it */
        const bool use = accept && nd.body != self;
        double h0 = 0.0, h1 = 0.0, h2 = 0.0, gm = 0.0;
        double c0 = 0.0, c1 = 0.0, c2 = 0.0, c3 = 0.0, c4 = 0.0, c5 = 0.0;
        if (use) {
            const double s2 = drSq + eps2;
            const double rinv = 1.0 / sqrt(s2);
            const double s3 = rinv / s2;
            const double s5 = s3 / s2;
            const double dg = dx * g0 + dy * g1 + dz * g2;
            double cg = nd.m * s3;                 /* coefficient of g */
            double cd = -3.0 * nd.m * s5 * dg;     /* coefficient of d */
            gm = s3 * dg;
            if (QUAD && cell) {
                const double* Q = nquad + 6 * q;
                const double qxx = Q[0], qxy = Q[1], qxz = Q[2];
                const double qyy = Q[3], qyz = Q[4], qzz = Q[5];
                const double ux = qxx * dx + qxy * dy + qxz * dz;
                const double uy = qxy * dx + qyy * dy + qyz * dz;
                const double uz = qxz * dx + qyz * dy + qzz * dz;
                const double vx = qxx * g0 + qxy * g1 + qxz * g2;
                const double vy = qxy * g0 + qyy * g1 + qyz * g2;
                const double vz = qxz * g0 + qyz * g1 + qzz * g2;
                const double dQd = ux * dx + uy * dy + uz * dz;
                const double ug = ux * g0 + uy * g1 + uz * g2;
                const double s7 = s5 / s2;
                const double s9 = s7 / s2;
                cg += 2.5 * dQd * s7;
                cd += -17.5 * dQd * s9 * dg + 5.0 * s7 * ug;
                const double cu = 5.0 * s7 * dg;
                h0 = cg * g0 + cd * dx + cu * ux - s5 * vx;
                h1 = cg * g1 + cd * dy + cu * uy - s5 * vy;
                h2 = cg * g2 + cd * dz + cu * uz - s5 * vz;
                const double a = 2.5 * s7 * dg;
                c0 = a * dx * dx - s5 * dx * g0;
                c1 = 2.0 * a * dx * dy - s5 * (dy * g0 + dx * g1);
                c2 = 2.0 * a * dx * dz - s5 * (dz * g0 + dx * g2);
                c3 = a * dy * dy - s5 * dy * g1;
                c4 = 2.0 * a * dy * dz - s5 * (dz * g1 + dy * g2);
                c5 = a * dz * dz - s5 * dz * g2;
            } else {
                h0 = cg * g0 + cd * dx;
                h1 = cg * g1 + cd * dy;
                h2 = cg * g2 + cd * dz;
            }
            ox -= h0;
            oy -= h1;
            oz -= h2;
        }
        if (accept) lp = nd_next;
        else if (open) lp = q + 1;
        if (o.multi && lp >= my_end) lp = INT_MAX;
        if (__builtin_amdgcn_ballot_w64(use) != 0ull) {
            h0 = nbh_wave_sum_f64(h0);
            h1 = nbh_wave_sum_f64(h1);
            h2 = nbh_wave_sum_f64(h2);
            gm = nbh_wave_sum_f64(gm);
            if (QUAD && cell) {
                c0 = nbh_wave_sum_f64(c0);
                c1 = nbh_wave_sum_f64(c1);
                c2 = nbh_wave_sum_f64(c2);
                c3 = nbh_wave_sum_f64(c3);
                c4 = nbh_wave_sum_f64(c4);
                c5 = nbh_wave_sum_f64(c5);
            }
            if (__lane_id() == 0) {
                unsafeAtomicAdd(&gnode[q], h0);
                unsafeAtomicAdd(&gnode[cap + q], h1);
                unsafeAtomicAdd(&gnode[2 * cap + q], h2);
                unsafeAtomicAdd(&gnode[3 * cap + q], gm);
                if (QUAD && cell) {
                    unsafeAtomicAdd(&gnode[4 * cap + q], c0);
                    unsafeAtomicAdd(&gnode[5 * cap + q], c1);
                    unsafeAtomicAdd(&gnode[6 * cap + q], c2);
                    unsafeAtomicAdd(&gnode[7 * cap + q], c3);
                    unsafeAtomicAdd(&gnode[8 * cap + q], c4);
                    unsafeAtomicAdd(&gnode[9 * cap + q], c5);
                }
            }
        }
        q = qn;
    }
    if (!valid) return;
    if (o.use_pot) {
        double ex, ey, ez;
        nbh_ext_jac(o.pot, p.x, p.y, p.z, g0, g1, g2, ex, ey, ez);
        ox += ex;
        oy += ey;
        oz += ez;
    }
    gxx[j] += ox;
    gxy[j] += oy;
    gxz[j] += oz;
}

/* Cells of one level hand their adjoint to their children (each child has
 * one parent, so no atomics). */
__global__ __launch_bounds__(256) void k_adj_down(const int* __restrict__ level_list, int level,
                                                  const NodeW* __restrict__ node,
                                                  double* __restrict__ gnode, int cap, int quad,
                                                  const DevScalars* __restrict__ scal)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= scal->level_count[level]) return;
    const int p = level_list[scal->level_off[level] + i];
    const NodeW c = node[p];
    const double Mp = c.m;
    const double Gx = gnode[p], Gy = gnode[cap + p], Gz = gnode[2 * cap + p];
    const double GM = gnode[3 * cap + p];
    double G[6] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
    if (quad)
        for (int k = 0; k < 6; ++k) G[k] = gnode[(4 + k) * cap + p];
    for (int ch = p + 1; ch < c.next;) {
        const NodeW d = node[ch];
        const double dx = d.x - c.x, dy = d.y - c.y, dz = d.z - c.z;
        const double w = d.m / Mp;
        double ax = w * Gx, ay = w * Gy, az = w * Gz;
        double am = GM + (Gx * dx + Gy * dy + Gz * dz) / Mp;
        if (quad) {
            const double m = d.m;
            ax += m * (4.0 * dx * G[0] + 3.0 * dy * G[1] + 3.0 * dz * G[2] - 2.0 * dx * G[3] -
                       2.0 * dx * G[5]);
            ay += m * (-2.0 * dy * G[0] + 3.0 * dx * G[1] + 4.0 * dy * G[3] + 3.0 * dz * G[4] -
                       2.0 * dy * G[5]);
            az += m * (-2.0 * dz * G[0] + 3.0 * dx * G[2] - 2.0 * dz * G[3] + 3.0 * dy * G[4] +
                       4.0 * dz * G[5]);
            const double r2 = dx * dx + dy * dy + dz * dz;
            am += G[0] * (3.0 * dx * dx - r2) + G[1] * 3.0 * dx * dy + G[2] * 3.0 * dx * dz +
                  G[3] * (3.0 * dy * dy - r2) + G[4] * 3.0 * dy * dz + G[5] * (3.0 * dz * dz - r2);
            if (d.body < 0)
                for (int k = 0; k < 6; ++k) gnode[(4 + k) * cap + ch] += G[k];
        }
        gnode[ch] += ax;
        gnode[cap + ch] += ay;
        gnode[2 * cap + ch] += az;
        gnode[3 * cap + ch] += am;
        ch = d.next;
    }
}

/* Adds v to acc[sim], summing across the wave first when every lane of the
 * wave belongs to one simulation. */
__device__ __forceinline__ void nbh_sim_add(double* acc, int sim, bool valid, double v)
{
    if (!valid) v = 0.0;
    const int sim0 = __builtin_amdgcn_readfirstlane(valid ? sim : 0x7fffffff);
    const bool same = __ballot(valid && sim != sim0) == 0ull;
    if (same) {
        v = nbh_wave_sum_f64(v);
        if (__lane_id() == 0 && sim0 != 0x7fffffff) unsafeAtomicAdd(&acc[sim0], v);
    } else if (valid) {
        unsafeAtomicAdd(&acc[sim], v);
    }
}

/* Adjoint of the closing half-kick v' = v_h + dt/2 a' (first half of the
 * reverse step): ga += dt/2 gv, dL/ddt += a'.gv / 2. */
__global__ __launch_bounds__(256) void k_adj_kick_end(int n, double c, const double* __restrict__ g,
                                                      double* __restrict__ ga,
                                                      const double* __restrict__ a,
                                                      const int* __restrict__ simid,
                                                      double* __restrict__ gdt)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < n;
    double contrib = 0.0;
    int sim = 0;
    if (valid) {
        const double* gv = g + 3 * (size_t) n;
        double* gax = ga;
        const double gvx = gv[i], gvy = gv[(size_t) n + i], gvz = gv[2 * (size_t) n + i];
        gax[i] += c * gvx;
        gax[(size_t) n + i] += c * gvy;
        gax[2 * (size_t) n + i] += c * gvz;
        contrib = 0.5 * (a[i] * gvx + a[(size_t) n + i] * gvy + a[2 * (size_t) n + i] * gvz);
        sim = simid[i];
    }
    nbh_sim_add(gdt, sim, valid, contrib);
}

/* Leaf records to bodies; with drift_kick, the adjoint of the drift
 * x' = x + dt v_h and of the opening half-kick v_h = v + dt/2 a:
 *   gv_h = gv + dt gx,  dL/ddt += v_h.gx + a.gv_h / 2,  gv = gv_h,  ga = dt/2 gv_h. */
__global__ __launch_bounds__(256) void k_adj_body(int n, int tree, const unsigned* __restrict__ perm,
                                                  const unsigned* __restrict__ cincl,
                                                  const double* __restrict__ gnode, int cap,
                                                  double* __restrict__ g, int drift_kick,
                                                  double dt, double c,
                                                  const double* __restrict__ vprev,
                                                  const double* __restrict__ aprev,
                                                  const int* __restrict__ simid,
                                                  double* __restrict__ gdt,
                                                  const DevScalars* __restrict__ scal)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    const bool valid = s < n;
    const size_t N = (size_t) n;
    double* gx = g;
    double* gv = g + 3 * N;
    double* ga = g + 6 * N;
    double* gm = g + 9 * N;
    int j = s, sim = 0;
    double contrib = 0.0;
    if (valid) {
        if (tree) {
            j = (int) perm[s];
            if (s < scal->nmassive) {
                const int leaf = s + (int) cincl[s];
                gx[j] += gnode[leaf];
                gx[N + j] += gnode[cap + leaf];
                gx[2 * N + j] += gnode[2 * cap + leaf];
                gm[j] += gnode[3 * cap + leaf];
            }
        }
        if (drift_kick) {
            double gvh[3], vh[3], ap[3];
            for (int d = 0; d < 3; ++d) {
                ap[d] = aprev[d * N + j];
                vh[d] = vprev[d * N + j] + c * ap[d];
                gvh[d] = gv[d * N + j] + dt * gx[d * N + j];
            }
            for (int d = 0; d < 3; ++d) {
                contrib += vh[d] * gx[d * N + j] + 0.5 * ap[d] * gvh[d];
                gv[d * N + j] = gvh[d];
                ga[d * N + j] = c * gvh[d];
            }
            sim = simid[j];
        }
    }
    if (drift_kick) nbh_sim_add(gdt, sim, valid, contrib);
}

static inline int nbh_blocks(int n, int per) { return n > 0 ? (n + per - 1) / per : 1; }

void nbh_adj_kick_end(nbhip_state* st, const double* a_next)
{
    NbhAdjoint* A = st->adj;
    const size_t N = (size_t) st->n;
    hipLaunchKernelGGL(k_adj_kick_end, dim3(nbh_blocks(st->n, 256)), dim3(256), 0, st->stream,
                       st->n, 0.5 * st->ctx.timestep, A->g, A->g + 6 * N, a_next, st->simid,
                       A->gdt);
}

/* Force adjoint at the current positions: the tree (built for them) or the
 * exact sum; reads ga, adds into gx and gm (tree: through gnode, finished by
 * nbh_adj_body). */
void nbh_adj_force(nbhip_state* st)
{
    NbhAdjoint* A = st->adj;
    const nbhip_ctx& c = st->ctx;
    hipStream_t s = st->stream;
    const size_t N = (size_t) st->n;
    if (c.criterion == NBHIP_EXACT) {
        nbh_exact_adjoint(st, A->g);
        return;
    }
    const int cap = st->node_cap;
    (void) hipMemsetAsync(A->gnode, 0, 10 * (size_t) cap * sizeof(double), s);
    AdjWalk o;
    o.n = st->n;
    o.use_pot = c.use_potential;
    o.multi = st->nsim > 1;
    o.cap = cap;
    o.pot = c.pot;
    const double* ga = A->g + 6 * N;
    dim3 grid(nbh_blocks(st->n, 256)), block(256);
    if (c.use_quad)
        hipLaunchKernelGGL(k_adj_walk<true>, grid, block, 0, s, st->node, st->nquad, st->spos,
                           st->perm, st->simid_s, st->sims, ga, ga + N, ga + 2 * N, A->g,
                           A->g + N, A->g + 2 * N, A->gnode, st->scal, o);
    else
        hipLaunchKernelGGL(k_adj_walk<false>, grid, block, 0, s, st->node, st->nquad, st->spos,
                           st->perm, st->simid_s, st->sims, ga, ga + N, ga + 2 * N, A->g,
                           A->g + N, A->g + 2 * N, A->gnode, st->scal, o);
    /* roots first: a cell's record is complete before it is handed down */
    if (!st->h_levels_valid && nbh_level_lists(st) != NBHIP_SUCCESS) return;
    for (int l = 0; l < 64; ++l) {
        int cnt = st->h_levels_valid ? st->h_level_count[l] : st->node_cap;
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(k_adj_down, dim3(nbh_blocks(cnt, 256)), dim3(256), 0, s,
                           st->level_list, l, st->node, A->gnode, cap, c.use_quad, st->scal);
    }
}

void nbh_adj_body(nbhip_state* st, int drift_kick, const double* v_prev, const double* a_prev)
{
    NbhAdjoint* A = st->adj;
    const int tree = st->ctx.criterion != NBHIP_EXACT;
    hipLaunchKernelGGL(k_adj_body, dim3(nbh_blocks(st->n, 256)), dim3(256), 0, st->stream, st->n,
                       tree, st->perm, st->cincl, A->gnode, st->node_cap, A->g, drift_kick,
                       st->ctx.timestep, 0.5 * st->ctx.timestep, v_prev, a_prev, st->simid,
                       A->gdt, st->scal);
}
