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
#include "nbhip_walk.h"

struct Quad6 {
    double q[6];
};

/* Plain form: FP64 end to end, reference operation order. */
template <int BS, bool QUAD, bool COUNT>
__global__ __launch_bounds__(BS) void k_walk64(const NodeW* __restrict__ node,
                                               const double* __restrict__ nquad,
                                               const double4* __restrict__ spos,
                                               const unsigned* __restrict__ perm,
                                               const int* __restrict__ simid_s,
                                               SimTab* __restrict__ sims,
                                               double* __restrict__ oax,
                                               double* __restrict__ oay,
                                               double* __restrict__ oaz,
                                               double* __restrict__ vx,
                                               double* __restrict__ vy,
                                               double* __restrict__ vz,
                                               DevScalars* __restrict__ scal, WalkOut o)
{
    const int s = blockIdx.x * BS + threadIdx.x;
    if ((scal->status & (NBHIP_CELL_OVERFLOW_ERROR | NBHIP_ERROR)) || scal->halt) return;
    const int M = scal->nmassive;
    const bool valid = s < o.n;
    if (__ballot(valid) == 0ull) return;
    const double4 p = valid ? spos[s] : make_double4(0.0, 0.0, 0.0, 0.0);
    const LaneStart L = nbh_lane_start(s, valid, M, simid_s, sims);
    const int self = L.self, my_end = L.end;
    int lp = L.lp;
    const double eps2 = L.eps2;
    const bool eps2_pos = o.all_eps2_pos;
    double ax = 0.0, ay = 0.0, az = 0.0;
    bool met = false;
    unsigned nbb = 0, nbc = 0, nop = 0, steps = 0;

    /* The opening tests come first: once the wave's vote is known the
     * successor node is fetched, and its scalar-load latency overlaps the
     * interaction math of the current node. */
    bool bad = false;
    int q = L.q0;
    NodeW nd = node[max(0, min(q, scal->nnode - 1))];
    /* the node's quadrupole moments ride along with its record (a cell's;
     * a body's slot is never read) */
    Quad6 nq = *reinterpret_cast<const Quad6*>(nquad + 6 * (size_t) max(0, min(q, scal->nnode - 1)));
    while (q < L.qend) {
        if (COUNT) ++steps;
        const int nd_next = __builtin_amdgcn_readfirstlane(nd.next);
        const int nd_body = __builtin_amdgcn_readfirstlane(nd.body);
        const bool active = lp == q;
        const double dx = nd.x - p.x;
        const double dy = nd.y - p.y;
        const double dz = nd.z - p.z;
        double drSq = dz * dz + (dy * dy + dx * dx);
        const bool accept = active && (nd_body >= 0 || drSq >= nd.rcrit2);
        const bool open = active && !accept;
        const bool any_open = __builtin_amdgcn_ballot_w64(open) != 0ull;
        int qn = any_open ? q + 1 : nd_next;
        if (qn <= q) {
            /* a Next() link that does not advance: corrupt table, stop */
            bad = true;
            qn = L.qend;
        }
        const NodeW nxt = node[min(qn, L.qend - 1)];
        const Quad6 nqx = *reinterpret_cast<const Quad6*>(nquad + 6 * (size_t) min(qn, L.qend - 1));
        if (accept) {
            /* with eps2 > 0 the self term adds exactly +0 (dx = dy = dz = 0),
             * so lanes need not branch around it */
            if (eps2_pos || nd_body != self) {
                drSq += eps2;
                const double drab = sqrt(drSq);
                const double phii = nd.m / drab;
                const double mor3 = phii / drSq;
                ax += mor3 * dx;
                ay += mor3 * dy;
                az += mor3 * dz;
                if (QUAD && nd_body < 0) {
                    const double qxx = nq.q[0], qxy = nq.q[1], qxz = nq.q[2];
                    const double qyy = nq.q[3], qyz = nq.q[4], qzz = nq.q[5];
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
                    if (nd_body >= 0) nbb += nd_body != self; else ++nbc;
                }
            }
            met |= nd_body == self;
            lp = nd_next;
        } else if (open) {
            lp = q + 1;
            if (COUNT) ++nop;
        }
        if (o.multi && lp >= my_end) lp = INT_MAX;
        q = qn;
        nd = nxt;
        nq = nqx;
    }

    if (bad && __lane_id() == 0) atomicOr(&scal->status, NBHIP_ERROR);
    if (COUNT) nbh_walk_counters(scal, nbb, nbc, nop, steps);
    if (!valid) return;
    if (!met) atomicOr(&sims[L.sim].incest, 1u);
    nbh_walk_finish(o, p.x, p.y, p.z, ax, ay, az, perm[s], oax, oay, oaz, vx, vy, vz);
}

/* Mixed form: displacements and interactions in FP32; FP32 partial sums
 * per block of 2^batch_shift pre-order node indices (counted from the
 * body's own simulation root), added into FP64 with Kahan compensation.
 * Positions, potential and kicks stay FP64.
 *
 * The walk is two nested loops: the inner one streams nodes until the wave
 * reaches the nearest block boundary of any of its lanes (q only grows), the
 * outer one flushes the lanes whose own block changed.  A lane's partial sum
 * therefore covers exactly its interactions inside one of its blocks -- a
 * property of its own interaction list, not of the wave -- and the FP64
 * state is not touched in the inner loop.
 *
 * COH selects the opening rule: false = each lane's own reference decision
 * (the FP64 walk's interaction lists); true = wave-coherent, the wave opens
 * a node when any lane's criterion says open and otherwise every lane of the
 * node's simulation interacts with it (each lane's list then refines its
 * reference list: every accepted node passes that lane's own test). */

template <int BS, bool QUAD, bool COUNT, bool COH, bool MULTI, bool EPS0>
__global__ __launch_bounds__(BS) void k_walk32(const NodeF* __restrict__ node,
                                               const NodeW* __restrict__ nodew,
                                               const double4* __restrict__ spos,
                                               const unsigned* __restrict__ perm,
                                               const int* __restrict__ simid_s,
                                               SimTab* __restrict__ sims,
                                               double* __restrict__ oax,
                                               double* __restrict__ oay,
                                               double* __restrict__ oaz,
                                               double* __restrict__ vx,
                                               double* __restrict__ vy,
                                               double* __restrict__ vz,
                                               DevScalars* __restrict__ scal, WalkOut o)
{
    const int s = blockIdx.x * BS + threadIdx.x;
    if ((scal->status & (NBHIP_CELL_OVERFLOW_ERROR | NBHIP_ERROR)) || scal->halt) return;
    const int M = scal->nmassive;
    const bool valid = s < o.n;
    if (__ballot(valid) == 0ull) return;
    const double4 p = valid ? spos[s] : make_double4(0.0, 0.0, 0.0, 0.0);
    const LaneStart L = nbh_lane_start(s, valid, M, simid_s, sims);
    const int self = L.self, my_end = L.end, my_root = L.lp;
    int lp = L.lp;
    const int root = sims[L.sim].root;
    const float eps2 = (float) L.eps2;
    const int bshift = o.batch_shift;
    const int bmask = (1 << bshift) - 1;
    const int qend = L.qend;
    /* every lane of the wave in one simulation: one root, one walk segment */
    const int root_u = __builtin_amdgcn_readfirstlane(root);
    const bool one_root = !MULTI ||
                          __builtin_amdgcn_ballot_w64(valid && root != root_u) == 0ull;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    double ax = 0.0, ay = 0.0, az = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    unsigned metv = ~0u;   /* 0 once the body met its own record */
    unsigned nbb = 0, nbc = 0, nop = 0, steps = 0;
    int q = L.q0;
    const unsigned long long valid_m = __builtin_amdgcn_ballot_w64(valid);

    /* One node step: test, vote, successor (fetched now, used next step),
     * interaction.  The successor lies in [q + 1, qend], so every step
     * advances and every fetch stays inside the table (it holds one record
     * past the last node); the summaries already reject broken Next links. */
    auto step = [&](const NodeF& nd, NodeF& nxt) {
        if (COUNT) ++steps;
        const float dx = (float) (nd.x - p.x);
        const float dy = (float) (nd.y - p.y);
        const float dz = (float) (nd.z - p.z);
        const float drSq = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
        bool accept, any_open;
        if (COH) {
            const bool active = one_root ? valid : (q >= my_root && q < my_end);
            const unsigned long long act = one_root ? valid_m : __builtin_amdgcn_ballot_w64(active);
            /* the lanes whose criterion accepts, straight from the comparison's
             * lane mask (ordered >=, so NaN opens); a ballot of the lane bool
             * would round-trip it through a vector register */
            const unsigned long long pass = __builtin_amdgcn_fcmpf(drSq, nd.rcrit2, 3);
            any_open = (act & ~pass) != 0ull;
            accept = active && !any_open;
        } else {
            const bool active = lp == q;
            accept = active && drSq >= nd.rcrit2;
            const bool opn = active && !accept;
            any_open = __builtin_amdgcn_ballot_w64(opn) != 0ull;
            if (COUNT) nop += opn;
            lp = accept ? nd.next : (opn ? q + 1 : lp);
            if (MULTI && lp >= my_end) lp = INT_MAX;
        }
        const int qn = min(max(any_open ? q + 1 : nd.next, q + 1), qend);
        nxt = node[qn];
        if (accept) {
            metv = min(metv, (unsigned) (nd.body ^ self));
            /* with eps2 = 0 the self term is skipped: d = 0 against 1 gives 0 */
            const float d2 = (EPS0 && nd.body == self) ? 1.0f : drSq + eps2;
            const float rinv = __builtin_amdgcn_rsqf(d2);
            const float rinv2 = rinv * rinv;
            if (QUAD) {
                /* acc += r^-5 ((m r^2 + 2.5 dQd / r^2) d - Qd); a body's
                 * record carries Q = 0, which leaves the monopole */
                const float qx = __builtin_fmaf(nd.q[0], dx, __builtin_fmaf(nd.q[1], dy, nd.q[2] * dz));
                const float qy = __builtin_fmaf(nd.q[1], dx, __builtin_fmaf(nd.q[3], dy, nd.q[4] * dz));
                const float qz = __builtin_fmaf(nd.q[2], dx, __builtin_fmaf(nd.q[4], dy, nd.q[5] * dz));
                const float dQd = __builtin_fmaf(qx, dx, __builtin_fmaf(qy, dy, qz * dz));
                const float dr5 = rinv2 * rinv2 * rinv;
                const float S = __builtin_fmaf(dQd, 2.5f * rinv2, nd.m * d2);
                sx = __builtin_fmaf(dr5, __builtin_fmaf(S, dx, -qx), sx);
                sy = __builtin_fmaf(dr5, __builtin_fmaf(S, dy, -qy), sy);
                sz = __builtin_fmaf(dr5, __builtin_fmaf(S, dz, -qz), sz);
            } else {
                const float mor3 = nd.m * rinv * rinv2;
                sx = __builtin_fmaf(mor3, dx, sx);
                sy = __builtin_fmaf(mor3, dy, sy);
                sz = __builtin_fmaf(mor3, dz, sz);
            }
            if (COUNT) {
                if (nd.body >= 0) nbb += nd.body != self; else ++nbc;
            }
        }
        q = qn;
    };

    NodeF nda = node[min(max(q, 0), qend)];
    NodeF ndb;
    while (q < qend) {
        /* nearest block boundary above q among the lanes' own blocks */
        int lim;
        if (one_root) {
            lim = root_u + ((q - root_u) | bmask) + 1;
        } else {
            const int mine = (valid && root >= 0) ? root + ((q - root) | bmask) + 1 : INT_MAX;
            lim = __builtin_amdgcn_readfirstlane(nbh_wave_min_i(mine));
        }
        lim = min(lim, qend);
        const int qstart = q;
        /* two steps per trip with the node records alternating between two
         * register sets */
        for (;;) {
            step(nda, ndb);
            if (q >= lim) {
                nda = ndb;
                break;
            }
            step(ndb, nda);
            if (q >= lim) break;
        }
        /* lanes whose own block changed close their FP32 partial sum (a sum
         * that is exactly zero leaves the FP64 state alone: the rule depends
         * on the lane's own interactions only, and it needs no per-step flag) */
        const int rb = one_root ? root_u : root;
        const bool crossed = ((q - rb) >> bshift) != ((qstart - rb) >> bshift);
        if (crossed && (sx != 0.0f || sy != 0.0f || sz != 0.0f)) {
            nbh_kahan_add(ax, cx, (double) sx);
            nbh_kahan_add(ay, cy, (double) sy);
            nbh_kahan_add(az, cz, (double) sz);
            sx = sy = sz = 0.0f;
        }
    }
    nbh_kahan_add(ax, cx, (double) sx);
    nbh_kahan_add(ay, cy, (double) sy);
    nbh_kahan_add(az, cz, (double) sz);

    if (COUNT) nbh_walk_counters(scal, nbb, nbc, nop, steps);
    if (!valid) return;
    if (metv != 0u) atomicOr(&sims[L.sim].incest, 1u);
    nbh_walk_finish(o, p.x, p.y, p.z, ax, ay, az, perm[s], oax, oay, oaz, vx, vy, vz);
}

/* FP64 direct-sum checker on sampled bodies: one block per sample, the
 * sum over sources split across 256 lanes and reduced in a fixed order. */
__global__ __launch_bounds__(256) void k_direct_sample(const double* __restrict__ px,
                                                       const double* __restrict__ py,
                                                       const double* __restrict__ pz,
                                                       const double* __restrict__ m,
                                                       const int* __restrict__ simid,
                                                       const SimTab* __restrict__ sims,
                                                       const int* __restrict__ idx,
                                                       double* __restrict__ acc)
{
    const int b = idx[blockIdx.x];
    const double x0 = px[b], y0 = py[b], z0 = pz[b];
    const SimTab& t = sims[simid[b]];
    const double eps2 = t.eps2;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int k = t.body_off + (int) threadIdx.x; k < t.body_off + t.body_n; k += 256) {
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
              int self_only, const WalkTune* tune)
{
    const nbhip_ctx& c = st->ctx;
    hipStream_t s = st->stream;
    if (c.criterion == NBHIP_EXACT) {
        nbh_exact_force(st, kick, oax, oay, oaz, self_only ? 0 : c.use_potential);
        return;
    }
    const WalkTune t = tune ? *tune : st->tune;
    WalkOut o;
    o.n = st->n;
    o.kick = kick;
    o.use_pot = self_only ? 0 : c.use_potential;
    o.multi = st->nsim > 1;
    o.all_eps2_pos = st->all_eps2_pos;
    int bsh = 0;
    while (bsh < 30 && (1 << (bsh + 1)) <= (t.batch > 0 ? t.batch : 1024)) ++bsh;
    o.batch_shift = bsh;
    o.far2 = st->group_far * st->group_far;
    o.dtHalf = 0.5 * c.timestep;
    o.pot = c.pot;
    if (c.precision == NBHIP_PRECISION_MIXED_GROUP) {
        nbh_walk_group(st, o, oax, oay, oaz, t, count);
        return;
    }
    const int bs = (t.block == 64 || t.block == 128) ? t.block : 256;
    dim3 grid(nbh_blocks(st->n, bs)), block(bs);
    if (c.precision != NBHIP_PRECISION_FP64) {
        const bool coherent = c.precision == NBHIP_PRECISION_MIXED_COHERENT;
#define NBH_W32E(B, Q, C, H, MU, E0)                                                         \
    hipLaunchKernelGGL((k_walk32<B, Q, C, H, MU, E0>), grid, block, 0, s, st->nodef, st->node, \
                       st->spos, st->perm, st->simid_s, st->sims, oax, oay, oaz, st->vx,       \
                       st->vy, st->vz, st->scal, o)
#define NBH_W32M(B, Q, C, H, MU)                                                             \
    do { if (o.all_eps2_pos) NBH_W32E(B, Q, C, H, MU, false);                                 \
         else NBH_W32E(B, Q, C, H, MU, true); } while (0)
#define NBH_W32C(B, Q, C, H)                                                                 \
    do { if (o.multi) NBH_W32M(B, Q, C, H, true); else NBH_W32M(B, Q, C, H, false); } while (0)
#define NBH_W32(B, Q, C)                                                                     \
    do { if (coherent) NBH_W32C(B, Q, C, true); else NBH_W32C(B, Q, C, false); } while (0)
#define NBH_W32B(B)                                                                          \
    if (c.use_quad) { if (count) NBH_W32(B, true, true); else NBH_W32(B, true, false); }     \
    else { if (count) NBH_W32(B, false, true); else NBH_W32(B, false, false); }
        if (bs == 64) { NBH_W32B(64) } else if (bs == 128) { NBH_W32B(128) } else { NBH_W32B(256) }
#undef NBH_W32B
#undef NBH_W32
#undef NBH_W32C
#undef NBH_W32M
#undef NBH_W32E
        return;
    }
#define NBH_W64(B, Q, C)                                                                     \
    hipLaunchKernelGGL((k_walk64<B, Q, C>), grid, block, 0, s, st->node, st->nquad, st->spos, \
                       st->perm, st->simid_s, st->sims, oax, oay, oaz, st->vx, st->vy, st->vz, \
                       st->scal, o)
#define NBH_W64B(B)                                                                          \
    if (c.use_quad) { if (count) NBH_W64(B, true, true); else NBH_W64(B, true, false); }     \
    else { if (count) NBH_W64(B, false, true); else NBH_W64(B, false, false); }
    if (bs == 64) { NBH_W64B(64) } else if (bs == 128) { NBH_W64B(128) } else { NBH_W64B(256) }
#undef NBH_W64B
#undef NBH_W64
}

void nbh_direct(nbhip_state* st, int nsample, const int* d_idx, double* d_acc)
{
    if (nsample <= 0) return;
    hipLaunchKernelGGL(k_direct_sample, dim3(nsample), dim3(256), 0, st->stream, st->px, st->py,
                       st->pz, st->m, st->simid, st->sims, d_idx, d_acc);
}
