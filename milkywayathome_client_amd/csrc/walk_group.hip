/*
 * LDS-staged group walk (precision NBHIP_PRECISION_MIXED_GROUP) on gfx950.
 *
 * A wavefront owns 64 consecutive bodies in Morton order.  Where the order
 * jumps between distant cells (the weakest common-prefix link of a run lies
 * more than two levels above the run's median link, or at level 4 or nearer
 * the root: bodies that share no cell smaller than a sixteenth of the root;
 * applied three times, so at most eight parts per run) and where a batch changes simulation, the wave is cut into
 * sub-groups, walked one after another.  (A group spanning two sibling
 * cells has a box twice theirs and walks several times their lists: at N =
 * 10^6 such groups, cut at three levels only, made single waves as long as
 * the whole walk.)  For a sub-group:
 *
 *  1. its bounding box is reduced across the member lanes (FP64);
 *  2. the tree is walked breadth-first 64 nodes at a time: the frontier is a
 *     stack of child ranges in LDS (each cell's children occupy one block of
 *     the walk table, so an opened cell pushes one entry), and lane l takes
 *     child (l mod 8) of entry (l / 8);
 *  3. a node is accepted when the box's nearest point passes the context's
 *     opening test (squared distance >= the record's rcrit2, which is the
 *     reference criterion's value rounded up), which implies every member
 *     body's own test (nbGravity, nbody/src/nbody_grav.c:42-124): each
 *     body's interaction list refines its reference list.  Bodies are always
 *     accepted; opened cells push their children;
 *  4. accepted nodes are appended to two LDS lists (cells with quadrupole,
 *     bodies) as FP32 positions relative to the group's centre (one FP64
 *     subtraction rounded once);
 *  5. a full list is drained by a dense loop in which every member lane
 *     interacts with every entry (monopole + quadrupole with the reference's
 *     softened expressions, nbody_grav.c:53-99, in FP32 with v_rsq_f32); one
 *     drained list is one FP32 partial sum, added into FP64 with Kahan
 *     compensation;
 *  6. with a far-field ratio delta > 0 (nbhip_set_group_expansion), an
 *     accepted cell whose distance D from the group's centre satisfies
 *     u_max <= delta |D| (u_max: half the diagonal of the group's box) is not
 *     evaluated per body: each lane takes one such cell and adds its softened
 *     field g = grad Phi, Jacobian J = grad grad Phi (monopole + quadrupole)
 *     and third-derivative tensor K (monopole) at the centre into 19 sums,
 *     which are added across the wave once per group; every member body then
 *     adds g - J u + K[u,u] / 2 at its offset u (a Taylor expansion whose
 *     remainder is O((u/|D|)^3) of the cell's field).  Far cells are kept at
 *     the back of the cell list (near ones at the front).
 *
 * The tests and the table are FP64 and deterministic, so a body's list does
 * not depend on timing, and repeated runs are bit-identical.
 */
#include "nbhip_walk.h"

#define NBH_GW_FCAP 512    /* range-stack entries per wavefront */
#ifndef NBH_GW_SPLIT_T
#define NBH_GW_SPLIT_T 2   /* split when min link level < median link level - 2 */
#endif
#ifndef NBH_GW_SPLIT_ABS
#define NBH_GW_SPLIT_ABS 4 /* and always when min link level <= 4 (cells of rsize / 16) */
#endif
#ifndef NBH_GW_SPLIT_PASSES
#define NBH_GW_SPLIT_PASSES 3   /* splitting passes: at most 2^passes sub-groups */
#endif

__device__ __forceinline__ double nbh_uniform_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int) (unsigned) (b & 0xffffffffll));
    const int hi = __builtin_amdgcn_readfirstlane((int) (b >> 32));
    return __longlong_as_double(((long long) hi << 32) | (unsigned) lo);
}

__device__ __forceinline__ double nbh_wave_min_d(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
    return nbh_uniform_d(v);
}

__device__ __forceinline__ double nbh_wave_max_d(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return nbh_uniform_d(v);
}

__device__ __forceinline__ int nbh_wave_min_u(int v)
{
    return __builtin_amdgcn_readfirstlane(nbh_wave_min_i(v));
}

__device__ __forceinline__ unsigned nbh_mbcnt(unsigned long long m)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned) (m >> 32),
                                     __builtin_amdgcn_mbcnt_lo((unsigned) m, 0u));
}

/* Cuts of one wave: bit i set = a sub-group ends at lane i.  `lvl` is the
 * level of the common prefix of this lane's body and the next lane's (-1
 * where the next lane is in another simulation, massless, invalid or past
 * the wave).  Runs between hard cuts are split at their weakest link (first
 * minimum) when that link's level is more than NBH_GW_SPLIT_T below (nearer
 * the root than) the lower median of the run's link levels, or is at most
 * NBH_GW_SPLIT_ABS (a sparse run: a few bodies far apart, whose box would
 * open most of the tree); NBH_GW_SPLIT_PASSES times. */
__device__ __forceinline__ unsigned long long nbh_group_cuts(unsigned long long hard, int lvl,
                                                             int lane)
{
    unsigned long long cuts = hard;
    for (int pass = 0; pass < NBH_GW_SPLIT_PASSES; ++pass) {
        unsigned long long next = cuts;
        int a = 0;
        unsigned long long rest = cuts;
        while (rest) {
            const int e = __builtin_ctzll(rest);      /* run [a, e] */
            rest &= rest - 1;
            const int nl = e - a;                      /* links a .. e-1 */
            if (nl >= 1) {
                const bool in = lane >= a && lane < e;
                const bool ok = __builtin_amdgcn_ballot_w64(in && lvl >= 0) ==
                                (((~0ull) >> (64 - nl)) << a);
                if (ok) {
                    const int key = nbh_wave_min_u(in ? ((lvl << 6) | lane) : 0x7fffffff);
                    const int lmin = key >> 6, j = key & 63;
                    /* lower median: least L with #{lvl <= L} >= ceil(nl / 2) */
                    const int need = (nl + 1) >> 1;
                    int lo = 0, hi = 63;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        const int c = __popcll(__builtin_amdgcn_ballot_w64(in && lvl <= mid));
                        if (c >= need) hi = mid; else lo = mid + 1;
                    }
                    if (lmin < lo - NBH_GW_SPLIT_T || lmin <= NBH_GW_SPLIT_ABS) next |= 1ull << j;
                }
            }
            a = e + 1;
        }
        cuts = next;
    }
    return cuts;
}

/* The sub-group cuts of the wave that holds sorted body s: hard cuts where
 * the next lane is invalid, in another simulation or of the other kind
 * (massive / massless), then nbh_group_cuts on the common-prefix levels. */
__device__ __forceinline__ unsigned long long nbh_wave_cuts(int s, bool valid, bool massive,
                                                            int sim, unsigned long long valid_m,
                                                            unsigned long long massive_m,
                                                            const int* __restrict__ dadj, int sb,
                                                            int lane)
{
    const int sim_next = __shfl_down(sim, 1, 64);
    const bool massive_next = lane < 63 && ((massive_m >> (lane + 1)) & 1ull);
    const bool valid_next = lane < 63 && ((valid_m >> (lane + 1)) & 1ull);
    const bool hard = valid && (!valid_next || sim_next != sim || massive_next != massive);
    int lvl = -1;
    if (massive && massive_next && !hard) lvl = nbh_levels(dadj[s], sb);
    return nbh_group_cuts(__builtin_amdgcn_ballot_w64(hard) | (1ull << 63), lvl, lane);
}

/* Members of the sub-group that starts at the lowest lane of `remaining`. */
__device__ __forceinline__ unsigned long long nbh_next_subgroup(unsigned long long remaining,
                                                                unsigned long long cuts)
{
    const int lead = __builtin_ctzll(remaining);
    const int last = __builtin_ctzll(cuts & (~0ull << lead));
    return (last == 63 ? ~0ull : ((1ull << (last + 1)) - 1ull)) & (~0ull << lead) & remaining;
}

template <int WPB, int LC, bool QUAD, bool EPS0, bool COUNT, bool FAR>
__global__ __launch_bounds__(64 * WPB) void k_walk_group(const NodeG* __restrict__ T,
                                                         const double4* __restrict__ spos,
                                                         const unsigned* __restrict__ perm,
                                                         const int* __restrict__ simid_s,
                                                         const int* __restrict__ dadj,
                                                         SimTab* __restrict__ sims,
                                                         double* __restrict__ oax,
                                                         double* __restrict__ oay,
                                                         double* __restrict__ oaz,
                                                         double* __restrict__ vx,
                                                         double* __restrict__ vy,
                                                         double* __restrict__ vz,
                                                         DevScalars* __restrict__ scal,
                                                         const int2* __restrict__ gorder,
                                                         unsigned* __restrict__ bcost,
                                                         float* __restrict__ bbox,
                                                         WalkOut o)
{
    __shared__ int s_fr[WPB][NBH_GW_FCAP];
    __shared__ float4 s_c0[WPB][LC];     /* cells: x y z (relative) m */
    __shared__ float4 s_c1[WPB][LC];     /* cells: q xx xy xz yy */
    __shared__ float2 s_c2[WPB][LC];     /* cells: q yz zz */
    __shared__ float4 s_b[WPB][LC];      /* bodies: x y z (relative) m */

    if ((scal->status & (NBHIP_CELL_OVERFLOW_ERROR | NBHIP_ERROR)) || scal->halt) return;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    int* fr = s_fr[w];
    float4* c0 = s_c0[w];
    float4* c1 = s_c1[w];
    float2* c2 = s_c2[w];
    float4* bl = s_b[w];

    /* this wave's task, costliest first: lanes [lo, hi) of a group, whole
     * sub-groups (the group itself, or a share of a costly one).  The cuts
     * are those of the whole group, so a body's list does not depend on how
     * its group was shared out. */
    const int slot = blockIdx.x * WPB + w;
    if (slot >= gorder[0].x) return;
    const int2 task = gorder[1 + slot];
    const int lo = task.y & 255, hi = task.y >> 8;
    const int s = task.x * 64 + lane;
    const bool in_group = s < o.n;
    const unsigned long long group_m = __builtin_amdgcn_ballot_w64(in_group);
    const int M = scal->nmassive;
    const int sb = scal->sim_bits;
    /* A depth-first descent pushes at most 7 entries more than it pops per
     * level below the entry it starts from; the stack keeps that much room
     * for the deepest tree of the state, so a throttled pop (one entry) can
     * always proceed without overflowing. */
    const int slack = 7 * (min(scal->maxdepth, 63) + 2);
    const int wave_base = s - lane;
    const double4 p = in_group ? spos[s] : make_double4(0.0, 0.0, 0.0, 0.0);
    const int sim = in_group ? simid_s[s] : -1;
    const bool gmassive = in_group && s < M;
    const unsigned long long massive_m = __builtin_amdgcn_ballot_w64(gmassive);
    const unsigned long long cuts =
        nbh_wave_cuts(s, in_group, gmassive, sim, group_m, massive_m, dadj, sb, lane);
    const bool valid = in_group && lane >= lo && lane < hi;
    const unsigned long long valid_m = __builtin_amdgcn_ballot_w64(valid);
    if (valid_m == 0ull) return;

    double ax = 0.0, ay = 0.0, az = 0.0, kx = 0.0, ky = 0.0, kz = 0.0;
    unsigned nbb = 0, nbc = 0, nop = 0, iters = 0, nfar = 0;
    unsigned work = 0;                 /* dense-loop entries + 4 per traversal round */
    unsigned my_work = 0;              /* the work of this lane's sub-group */
    float my_box = 0.0f;               /* and its box half-diagonal */
    unsigned self_hits = 0;
    bool bad = false;
    unsigned long long remaining = valid_m;
    while (remaining) {
        const int lead = __builtin_ctzll(remaining);
        const unsigned long long mem = nbh_next_subgroup(remaining, cuts);
        remaining &= ~mem;
        const unsigned work0 = work;
        const bool me = (mem >> lane) & 1ull;
        const int sim_u = __builtin_amdgcn_readlane(sim, lead);
        const int root = sims[sim_u].root;
        if (root < 0) continue;                       /* no massive body: no tree */
        const float eps2 = (float) sims[sim_u].eps2;

        /* bounding box and centre of the sub-group */
        const double lox = nbh_wave_min_d(me ? p.x : INFINITY);
        const double loy = nbh_wave_min_d(me ? p.y : INFINITY);
        const double loz = nbh_wave_min_d(me ? p.z : INFINITY);
        const double hix = nbh_wave_max_d(me ? p.x : -INFINITY);
        const double hiy = nbh_wave_max_d(me ? p.y : -INFINITY);
        const double hiz = nbh_wave_max_d(me ? p.z : -INFINITY);
        const double gx = 0.5 * (lox + hix), gy = 0.5 * (loy + hiy), gz = 0.5 * (loz + hiz);
        const float rx = me ? (float) (p.x - gx) : 0.0f;
        const float ry = me ? (float) (p.y - gy) : 0.0f;
        const float rz = me ? (float) (p.z - gz) : 0.0f;
        const int expected = __popcll(mem & massive_m);
        int own_seen = 0;
        /* far field: u_max^2 (half the box diagonal, squared) and the
         * lane's partial sums of g (3), J (xx xy xz yy yz zz) and K (xxx
         * xxy xxz xyy xyz xzz yyy yyz yzz zzz) */
        const double hbx = 0.5 * (hix - lox), hby = 0.5 * (hiy - loy), hbz = 0.5 * (hiz - loz);
        const double umax2 = hbz * hbz + (hby * hby + hbx * hbx);
        const bool use_far = FAR;
        float fg[3] = { 0.0f, 0.0f, 0.0f };
        float fj[6] = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };
        float fk[10] = { 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f };

        int fsize = 1;
        if (lane == 0) fr[0] = (sim_u << 3);          /* slot sim_u, one node */
        __builtin_amdgcn_wave_barrier();
        int nc = 0, nb = 0, nf = 0;

        /* drains of the two lists */
        auto flush_cells = [&](int cnt) {
            float tx = 0.0f, ty = 0.0f, tz = 0.0f;
            __builtin_amdgcn_wave_barrier();
            #pragma unroll 2
            for (int j = 0; j < cnt; ++j) {
                const float4 a = c0[j];
                const float dx = a.x - rx, dy = a.y - ry, dz = a.z - rz;
                float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
                if (EPS0) d2 = d2 == 0.0f ? 1.0f : d2;
                const float rinv = __builtin_amdgcn_rsqf(d2);
                const float rinv2 = rinv * rinv;
                if (QUAD) {
                    const float4 b = c1[j];
                    const float2 e = c2[j];
                    /* acc += r^-5 ((m r^2 + 2.5 dQd / r^2) d - Qd) */
                    const float qx = __builtin_fmaf(b.x, dx, __builtin_fmaf(b.y, dy, b.z * dz));
                    const float qy = __builtin_fmaf(b.y, dx, __builtin_fmaf(b.w, dy, e.x * dz));
                    const float qz = __builtin_fmaf(b.z, dx, __builtin_fmaf(e.x, dy, e.y * dz));
                    const float dQd = __builtin_fmaf(qx, dx, __builtin_fmaf(qy, dy, qz * dz));
                    const float dr5 = rinv2 * rinv2 * rinv;
                    const float S = __builtin_fmaf(dQd, 2.5f * rinv2, a.w * d2);
                    tx = __builtin_fmaf(dr5, __builtin_fmaf(S, dx, -qx), tx);
                    ty = __builtin_fmaf(dr5, __builtin_fmaf(S, dy, -qy), ty);
                    tz = __builtin_fmaf(dr5, __builtin_fmaf(S, dz, -qz), tz);
                } else {
                    const float mor3 = a.w * rinv * rinv2;
                    tx = __builtin_fmaf(mor3, dx, tx);
                    ty = __builtin_fmaf(mor3, dy, ty);
                    tz = __builtin_fmaf(mor3, dz, tz);
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (me) {
                nbh_kahan_add(ax, kx, (double) tx);
                nbh_kahan_add(ay, ky, (double) ty);
                nbh_kahan_add(az, kz, (double) tz);
            }
            if (COUNT) {
                nbc += me ? (unsigned) cnt : 0u;
                iters += (unsigned) cnt;
            }
            work += (unsigned) cnt;
        };
        /* far cells: slots LC-1, LC-2, ...; one cell per lane */
        auto flush_far = [&](int cnt) {
            __builtin_amdgcn_wave_barrier();
            for (int j0 = 0; j0 < cnt; j0 += 64) {
                const int e = j0 + lane;
                if (e < cnt) {
                    const int slot = LC - 1 - e;
                    const float4 a = c0[slot];
                    const float dx = a.x, dy = a.y, dz = a.z, m = a.w;
                    const float s2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy,
                                                    __builtin_fmaf(dz, dz, eps2)));
                    const float ri = __builtin_amdgcn_rsqf(s2);
                    const float ri2 = ri * ri;
                    const float ri3 = ri * ri2, ri5 = ri3 * ri2, ri7 = ri5 * ri2;
                    /* monopole: g = m d / s^3, J = m (I / s^3 - 3 d d^T / s^5),
                     * K = B (delta d + ...) + A d d d, B = -3 m / s^5, A = 15 m / s^7 */
                    const float mr3 = m * ri3;
                    const float B = -3.0f * m * ri5;
                    const float A = 15.0f * m * ri7;
                    const float xx = dx * dx, xy = dx * dy, xz = dx * dz;
                    const float yy = dy * dy, yz = dy * dz, zz = dz * dz;
                    fg[0] = __builtin_fmaf(mr3, dx, fg[0]);
                    fg[1] = __builtin_fmaf(mr3, dy, fg[1]);
                    fg[2] = __builtin_fmaf(mr3, dz, fg[2]);
                    fj[0] += __builtin_fmaf(B, xx, mr3);
                    fj[1] = __builtin_fmaf(B, xy, fj[1]);
                    fj[2] = __builtin_fmaf(B, xz, fj[2]);
                    fj[3] += __builtin_fmaf(B, yy, mr3);
                    fj[4] = __builtin_fmaf(B, yz, fj[4]);
                    fj[5] += __builtin_fmaf(B, zz, mr3);
                    const float Ax = A * dx, Ay = A * dy, Az = A * dz;
                    fk[0] += __builtin_fmaf(Ax, xx, 3.0f * B * dx);   /* xxx */
                    fk[1] += __builtin_fmaf(Ax, xy, B * dy);          /* xxy */
                    fk[2] += __builtin_fmaf(Ax, xz, B * dz);          /* xxz */
                    fk[3] += __builtin_fmaf(Ax, yy, B * dx);          /* xyy */
                    fk[4] = __builtin_fmaf(Ax, yz, fk[4]);            /* xyz */
                    fk[5] += __builtin_fmaf(Ax, zz, B * dx);          /* xzz */
                    fk[6] += __builtin_fmaf(Ay, yy, 3.0f * B * dy);   /* yyy */
                    fk[7] += __builtin_fmaf(Ay, yz, B * dz);          /* yyz */
                    fk[8] += __builtin_fmaf(Ay, zz, B * dy);          /* yzz */
                    fk[9] += __builtin_fmaf(Az, zz, 3.0f * B * dz);   /* zzz */
                    if (QUAD) {
                        /* g_Q = 2.5 (dQd) d / s^7 - Q d / s^5, J_Q = -Q / s^5
                         * + 5 (q d^T + d q^T) / s^7 + 2.5 dQd I / s^7
                         * - 17.5 dQd d d^T / s^9 (q = Q d) */
                        const float4 b = c1[slot];
                        const float2 e2 = c2[slot];
                        const float qx = __builtin_fmaf(b.x, dx, __builtin_fmaf(b.y, dy, b.z * dz));
                        const float qy = __builtin_fmaf(b.y, dx, __builtin_fmaf(b.w, dy, e2.x * dz));
                        const float qz = __builtin_fmaf(b.z, dx, __builtin_fmaf(e2.x, dy, e2.y * dz));
                        const float dqd = __builtin_fmaf(qx, dx, __builtin_fmaf(qy, dy, qz * dz));
                        const float h = 2.5f * dqd * ri7;
                        fg[0] += __builtin_fmaf(h, dx, -ri5 * qx);
                        fg[1] += __builtin_fmaf(h, dy, -ri5 * qy);
                        fg[2] += __builtin_fmaf(h, dz, -ri5 * qz);
                        const float f5 = 5.0f * ri7;
                        const float g9 = -17.5f * dqd * ri7 * ri2;
                        fj[0] += __builtin_fmaf(g9, xx, __builtin_fmaf(2.0f * f5, qx * dx,
                                                                       h - ri5 * b.x));
                        fj[1] += __builtin_fmaf(g9, xy, __builtin_fmaf(f5, __builtin_fmaf(qx, dy, qy * dx),
                                                                       -ri5 * b.y));
                        fj[2] += __builtin_fmaf(g9, xz, __builtin_fmaf(f5, __builtin_fmaf(qx, dz, qz * dx),
                                                                       -ri5 * b.z));
                        fj[3] += __builtin_fmaf(g9, yy, __builtin_fmaf(2.0f * f5, qy * dy,
                                                                       h - ri5 * b.w));
                        fj[4] += __builtin_fmaf(g9, yz, __builtin_fmaf(f5, __builtin_fmaf(qy, dz, qz * dy),
                                                                       -ri5 * e2.x));
                        fj[5] += __builtin_fmaf(g9, zz, __builtin_fmaf(2.0f * f5, qz * dz,
                                                                       h - ri5 * e2.y));
                    }
                }
            }
            __builtin_amdgcn_wave_barrier();
            if (COUNT) nfar += (unsigned) cnt;
        };
        auto flush_bodies = [&](int cnt) {
            float tx = 0.0f, ty = 0.0f, tz = 0.0f;
            __builtin_amdgcn_wave_barrier();
            #pragma unroll 2
            for (int j = 0; j < cnt; ++j) {
                const float4 a = bl[j];
                const float dx = a.x - rx, dy = a.y - ry, dz = a.z - rz;
                float d2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
                /* eps2 > 0: the body's own entry adds exactly 0 (d = 0);
                 * eps2 = 0: it is skipped by making d2 finite */
                if (EPS0) d2 = d2 == 0.0f ? 1.0f : d2;
                const float rinv = __builtin_amdgcn_rsqf(d2);
                const float mor3 = a.w * rinv * (rinv * rinv);
                tx = __builtin_fmaf(mor3, dx, tx);
                ty = __builtin_fmaf(mor3, dy, ty);
                tz = __builtin_fmaf(mor3, dz, tz);
            }
            __builtin_amdgcn_wave_barrier();
            if (me) {
                nbh_kahan_add(ax, kx, (double) tx);
                nbh_kahan_add(ay, ky, (double) ty);
                nbh_kahan_add(az, kz, (double) tz);
            }
            if (COUNT) {
                nbb += me ? (unsigned) cnt : 0u;
                iters += (unsigned) cnt;
            }
            work += (unsigned) cnt;
        };

        while (fsize > 0) {
            /* pop up to 8 ranges (64 nodes); throttled near the capacity so
             * that the pushes of this round always fit */
            int E = min(8, fsize);
            if (fsize + 7 * E > NBH_GW_FCAP - slack)
                E = max(1, min(E, (NBH_GW_FCAP - slack - fsize) / 7));
            const int eidx = lane >> 3, sub = lane & 7;
            const int ent = eidx < E ? fr[fsize - 1 - eidx] : 0;
            const int cnt = (ent & 7) + 1;
            const bool act = eidx < E && sub < cnt;
            fsize -= E;
            __builtin_amdgcn_wave_barrier();
            work += 4u;
            const NodeG r = T[act ? (ent >> 3) + sub : 0];
            /* FP64 test against the sub-group's box */
            const double ddx = fmax(fmax(lox - r.x, r.x - hix), 0.0);
            const double ddy = fmax(fmax(loy - r.y, r.y - hiy), 0.0);
            const double ddz = fmax(fmax(loz - r.z, r.z - hiz), 0.0);
            const double d2 = ddz * ddz + (ddy * ddy + ddx * ddx);
            const bool isbody = r.count == 0;
            const bool acc = act && (isbody || d2 >= (double) r.rcrit2);
            const bool opn = act && !acc;
            bool far = false;
            if (FAR) {
                const double Dx = r.x - gx, Dy = r.y - gy, Dz = r.z - gz;
                far = acc && !isbody && umax2 <= o.far2 * (Dz * Dz + (Dy * Dy + Dx * Dx));
            }
            const unsigned long long cm = __builtin_amdgcn_ballot_w64(acc && !isbody && !far);
            const unsigned long long fm = FAR ? __builtin_amdgcn_ballot_w64(far) : 0ull;
            const unsigned long long bm = __builtin_amdgcn_ballot_w64(acc && isbody);
            const unsigned long long om = __builtin_amdgcn_ballot_w64(opn);
            /* a member's own leaf: its body meets itself (no incest) */
            const unsigned rel = (unsigned) (r.link - wave_base);
            const bool own = acc && isbody && rel < 64u && ((mem >> rel) & 1ull);
            own_seen += __popcll(__builtin_amdgcn_ballot_w64(own));
            const int npush = __popcll(om);
            if (fsize + npush > NBH_GW_FCAP) {
                bad = true;                           /* deeper than 63 levels */
                break;
            }
            if (acc && !isbody) {
                const int slot = (FAR && far) ? LC - 1 - (nf + (int) nbh_mbcnt(fm))
                                     : nc + (int) nbh_mbcnt(cm);
                c0[slot] = make_float4((float) (r.x - gx), (float) (r.y - gy),
                                       (float) (r.z - gz), r.m);
                if (QUAD) {
                    c1[slot] = make_float4(r.q[0], r.q[1], r.q[2], r.q[3]);
                    c2[slot] = make_float2(r.q[4], r.q[5]);
                }
            }
            if (acc && isbody) {
                const int slot = nb + (int) nbh_mbcnt(bm);
                bl[slot] = make_float4((float) (r.x - gx), (float) (r.y - gy),
                                       (float) (r.z - gz), r.m);
            }
            if (opn) fr[fsize + (int) nbh_mbcnt(om)] = (r.link << 3) | (r.count - 1);
            __builtin_amdgcn_wave_barrier();
            nc += __popcll(cm);
            nf += __popcll(fm);
            nb += __popcll(bm);
            fsize += npush;
            if (COUNT) nop += (unsigned) npush;
            if (nc + nf > LC - 64) {
                if (nc > 0) flush_cells(nc);
                if (FAR && nf > 0) flush_far(nf);
                nc = 0;
                nf = 0;
            }
            if (nb > LC - 64) {
                flush_bodies(nb);
                nb = 0;
            }
        }
        if (nc > 0) flush_cells(nc);
        if (FAR && nf > 0) flush_far(nf);
        if (nb > 0) flush_bodies(nb);
        if (use_far) {
            /* the wave's sums (every lane holds the total), then each member
             * body adds g - J u + K[u,u] / 2 */
            for (int off = 32; off > 0; off >>= 1) {
                for (int k = 0; k < 3; ++k) fg[k] += __shfl_xor(fg[k], off, 64);
                for (int k = 0; k < 6; ++k) fj[k] += __shfl_xor(fj[k], off, 64);
                for (int k = 0; k < 10; ++k) fk[k] += __shfl_xor(fk[k], off, 64);
            }
            if (me) {
                const float ux = rx, uy = ry, uz = rz;
                const float uxx = 0.5f * ux * ux, uyy = 0.5f * uy * uy, uzz = 0.5f * uz * uz;
                const float uxy = ux * uy, uxz = ux * uz, uyz = uy * uz;
                float tx = fg[0] - __builtin_fmaf(fj[0], ux, __builtin_fmaf(fj[1], uy, fj[2] * uz));
                float ty = fg[1] - __builtin_fmaf(fj[1], ux, __builtin_fmaf(fj[3], uy, fj[4] * uz));
                float tz = fg[2] - __builtin_fmaf(fj[2], ux, __builtin_fmaf(fj[4], uy, fj[5] * uz));
                tx += fk[0] * uxx + fk[3] * uyy + fk[5] * uzz + fk[1] * uxy + fk[2] * uxz + fk[4] * uyz;
                ty += fk[1] * uxx + fk[6] * uyy + fk[8] * uzz + fk[3] * uxy + fk[4] * uxz + fk[7] * uyz;
                tz += fk[2] * uxx + fk[7] * uyy + fk[9] * uzz + fk[4] * uxy + fk[5] * uxz + fk[8] * uyz;
                nbh_kahan_add(ax, kx, (double) tx);
                nbh_kahan_add(ay, ky, (double) ty);
                nbh_kahan_add(az, kz, (double) tz);
            }
        }
        if (COUNT) self_hits += (unsigned) own_seen;
        if (own_seen < expected && lane == 0) atomicOr(&sims[sim_u].incest, 1u);
        if (me) {
            my_work = work - work0;
            my_box = (float) sqrt(umax2);
        }
        if (bad) break;
    }

    if (bad && lane == 0) atomicOr(&scal->status, NBHIP_ERROR);

    if (COUNT) {
        /* the self interactions (d = 0) are not counted, as in the reference;
         * nop and iters are wave-uniform */
        const unsigned long long c0 = nbh_wave_sum_u64(nbb);
        const unsigned long long c1 = nbh_wave_sum_u64(nbc);
        if (lane == 0) {
            atomicAdd(&scal->counters[0], c0 - self_hits);
            atomicAdd(&scal->counters[1], c1);
            atomicAdd(&scal->counters[2], (unsigned long long) nop);
            atomicAdd(&scal->counters[3], (unsigned long long) iters);
            atomicAdd(&scal->counters[4], (unsigned long long) nfar);
        }
    }
    if (!valid) return;
    const unsigned j = perm[s];
    bcost[j] = my_work;                /* the body carries its sub-group's cost to the next walk */
    bbox[j] = my_box;
    nbh_walk_finish(o, p.x, p.y, p.z, ax, ay, az, j, oax, oay, oaz, vx, vy, vz);
}

/* Predicted cost of each group (wave) and its tasks.  A sub-group (cut as
 * the walk cuts it) is priced at the mean over its bodies of each body's
 * last sub-group cost scaled by the ratio of the new box to that sub-group's
 * box (clamped to [1/4, 4]) -- the costs follow the bodies, a sub-group's
 * cost grows with its box, and a wave cut into parts walks each part -- and
 * a group at the sum over its sub-groups.  (At N = 10^6 this prediction
 * correlates at 0.98 with the walk's costs after a step, the largest last
 * cost of the members alone at 0.65.)  A group above
 * NBH_GW_SPLIT_COST times the last walk's mean group is shared out: runs of
 * whole sub-groups, a run closed before it would pass that cost, at most 8.
 * gcost[g] = min(cost, 2^28 - 1) | (tasks - 1) << 28; gtask[8 g + k] =
 * (cost, lo | hi << 8) of task k (lanes [lo, hi)). */
__global__ __launch_bounds__(256) void k_group_cost(const unsigned* __restrict__ bcost,
                                                    const float* __restrict__ bbox,
                                                    const double4* __restrict__ spos,
                                                    const unsigned* __restrict__ perm,
                                                    const int* __restrict__ simid_s,
                                                    const int* __restrict__ dadj,
                                                    const DevScalars* __restrict__ scal,
                                                    unsigned* __restrict__ gcost,
                                                    uint2* __restrict__ gtask, int n)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const int g = s >> 6;
    const bool valid = s < n;
    const unsigned long long valid_m = __builtin_amdgcn_ballot_w64(valid);
    if (valid_m == 0ull) return;
    const int M = scal->nmassive;
    const int sim = valid ? simid_s[s] : -1;
    const bool massive = valid && s < M;
    const unsigned long long massive_m = __builtin_amdgcn_ballot_w64(massive);
    const unsigned long long cuts =
        nbh_wave_cuts(s, valid, massive, sim, valid_m, massive_m, dadj, scal->sim_bits, lane);
    const unsigned j = valid ? perm[s] : 0u;
    const float c = valid ? (float) bcost[j] : 0.0f;
    const float bold = valid ? bbox[j] : 0.0f;
    const double4 p = valid ? spos[s] : make_double4(0.0, 0.0, 0.0, 0.0);
    const double limit = NBH_GW_SPLIT_COST * (double) scal->group_mean;
    unsigned total = 0, tcost = 0;
    int ntask = 0, tlo = 0;
    unsigned long long remaining = valid_m;
    while (remaining) {
        const unsigned long long mem = nbh_next_subgroup(remaining, cuts);
        remaining &= ~mem;
        const bool me = (mem >> lane) & 1ull;
        const double hx = 0.5 * (nbh_wave_max_d(me ? p.x : -INFINITY) - nbh_wave_min_d(me ? p.x : INFINITY));
        const double hy = 0.5 * (nbh_wave_max_d(me ? p.y : -INFINITY) - nbh_wave_min_d(me ? p.y : INFINITY));
        const double hz = 0.5 * (nbh_wave_max_d(me ? p.z : -INFINITY) - nbh_wave_min_d(me ? p.z : INFINITY));
        const float bnew = (float) sqrt(hz * hz + (hy * hy + hx * hx));
        float r = bold > 0.0f ? bnew / bold : (bnew > 0.0f ? 4.0f : 1.0f);
        r = fminf(fmaxf(r, 0.25f), 4.0f);
        float v = me ? c * r : 0.0f;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        const unsigned m = (unsigned) fminf(v / (float) __popcll(mem) + 0.5f, 268435455.0f);
        const int a = __builtin_ctzll(mem);
        /* wave-uniform from here */
        if (limit > 0.0 && tcost > 0u && (double) tcost + (double) m > limit && ntask < 7) {
            if (lane == 0) gtask[8 * (size_t) g + ntask] = make_uint2(tcost, (unsigned) (tlo | (a << 8)));
            ++ntask;
            tlo = a;
            tcost = 0u;
        }
        tcost += m;
        total += m;
    }
    if (lane == 0) {
        gtask[8 * (size_t) g + ntask] = make_uint2(tcost, (unsigned) (tlo | (64 << 8)));
        gcost[g] = min(total, (1u << 28) - 1u) | ((unsigned) ntask << 28);
    }
}

/* Tasks of the next group walk in dispatch order, costliest first (a
 * counting sort over 256 logarithmic cost buckets, one workgroup; the order
 * within a bucket is whatever the atomics give -- the results of the walk
 * depend neither on the order nor on the sharing, only its tail does).
 * gorder[0].x = the task count; entries (group, lo | hi << 8).  Also keeps
 * the mean group cost for the next walk's sharing. */
#define NBH_ORDER_THREADS 1024
__global__ __launch_bounds__(NBH_ORDER_THREADS) void k_group_order(unsigned* __restrict__ gcost,
                                                                  const uint2* __restrict__ gtask,
                                                                  int2* __restrict__ gorder,
                                                                  DevScalars* __restrict__ scal,
                                                                  int ngroup)
{
    __shared__ int cnt[256];
    __shared__ int off[256];
    __shared__ unsigned long long wsum[NBH_ORDER_THREADS / 64];
    __shared__ int extra;
    const int tid = threadIdx.x;
    if (tid < 256) cnt[tid] = 0;
    if (tid == 0) extra = 0;
    auto bucket = [](unsigned c) {
        const unsigned v = c + 1u;
        const int e = 31 - __clz(v);                         /* floor(log2 v), 0..31 */
        const int f = e >= 3 ? (int) ((v >> (e - 3)) & 7u) : (int) ((v << (3 - e)) & 7u);
        return 255 - min(255, e * 8 + f);                    /* descending cost */
    };
    const unsigned cmask = (1u << 28) - 1u;
    unsigned long long sum = 0;
    for (int g = tid; g < ngroup; g += NBH_ORDER_THREADS) sum += gcost[g] & cmask;
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
    if ((tid & 63) == 0) wsum[tid >> 6] = sum;
    __syncthreads();
    /* every task entry a group will have; shares past the extra-task budget
     * fall back to the whole group */
    for (int g = tid; g < ngroup; g += NBH_ORDER_THREADS) {
        const unsigned e = gcost[g];
        int k = (int) (e >> 28) + 1;
        if (k > 1 && atomicAdd(&extra, k - 1) + k - 1 > NBH_GW_EXTRA_TASKS) {
            k = 1;
            gcost[g] = e & cmask;
        }
        if (k == 1) atomicAdd(&cnt[bucket(e & cmask)], 1);
        else
            for (int t = 0; t < k; ++t) atomicAdd(&cnt[bucket(gtask[8 * (size_t) g + t].x)], 1);
    }
    __syncthreads();
    if (tid == 0) {
        unsigned long long total = 0;
        for (int k = 0; k < NBH_ORDER_THREADS / 64; ++k) total += wsum[k];
        scal->group_mean = (unsigned) min(total / (unsigned long long) max(ngroup, 1),
                                          (unsigned long long) cmask);
        int acc = 0;
        for (int b = 0; b < 256; ++b) {
            off[b] = acc;
            acc += cnt[b];
        }
        gorder[0] = make_int2(acc, 0);
    }
    __syncthreads();
    for (int g = tid; g < ngroup; g += NBH_ORDER_THREADS) {
        const unsigned e = gcost[g];
        const int k = (int) (e >> 28) + 1;
        if (k == 1) {
            gorder[1 + atomicAdd(&off[bucket(e & cmask)], 1)] = make_int2(g, 64 << 8);
        } else {
            for (int t = 0; t < k; ++t) {
                const uint2 tk = gtask[8 * (size_t) g + t];
                gorder[1 + atomicAdd(&off[bucket(tk.x)], 1)] = make_int2(g, (int) tk.y);
            }
        }
    }
}

static inline int nbh_gblocks(int n, int per) { return n > 0 ? (n + per - 1) / per : 1; }

void nbh_walk_group(nbhip_state* st, const WalkOut& o, double* oax, double* oay, double* oaz,
                    const WalkTune& t, int count)
{
    const int bs = (t.block == 64 || t.block == 128) ? t.block : 256;
    int lc = (t.batch == 64 || t.batch == 128) ? t.batch : 256;
    if (bs == 256 && lc > 128) lc = 128;       /* LDS: at most 64 KiB per workgroup */
    dim3 grid(nbh_gblocks(nbh_group_tasks(st->n) * 64, bs)), block(bs);
    hipLaunchKernelGGL(k_group_cost, dim3(nbh_gblocks(st->n, 256)), dim3(256), 0, st->stream,
                       st->bcost, st->bbox, st->spos, st->perm, st->simid_s, st->dadj, st->scal,
                       st->gcost, st->gtask, st->n);
    hipLaunchKernelGGL(k_group_order, dim3(1), dim3(NBH_ORDER_THREADS), 0, st->stream, st->gcost,
                       st->gtask, st->gorder, st->scal, (st->n + 63) / 64);
    const bool quad = st->ctx.use_quad != 0;
    const bool eps0 = !o.all_eps2_pos;
    const bool far = o.far2 > 0.0;
#define NBH_GWF(W, L, Q, E, C, F)                                                             \
    hipLaunchKernelGGL((k_walk_group<W, L, Q, E, C, F>), grid, block, 0, st->stream,           \
                       st->nodeg, st->spos, st->perm, st->simid_s, st->dadj, st->sims, oax,    \
                       oay, oaz, st->vx, st->vy, st->vz, st->scal, st->gorder, st->bcost, \
                       st->bbox, o)
#define NBH_GWL(W, L, Q, E, C)                                                                \
    do { if (far) NBH_GWF(W, L, Q, E, C, true); else NBH_GWF(W, L, Q, E, C, false); } while (0)
#define NBH_GWC(W, L, Q, E)                                                                   \
    do { if (count) NBH_GWL(W, L, Q, E, true); else NBH_GWL(W, L, Q, E, false); } while (0)
#define NBH_GWE(W, L, Q)                                                                      \
    do { if (eps0) NBH_GWC(W, L, Q, true); else NBH_GWC(W, L, Q, false); } while (0)
#define NBH_GWQ(W, L)                                                                         \
    do { if (quad) NBH_GWE(W, L, true); else NBH_GWE(W, L, false); } while (0)
    if (bs == 64) {
        if (lc == 64) NBH_GWQ(1, 64); else if (lc == 128) NBH_GWQ(1, 128); else NBH_GWQ(1, 256);
    } else if (bs == 128) {
        if (lc == 64) NBH_GWQ(2, 64); else if (lc == 128) NBH_GWQ(2, 128); else NBH_GWQ(2, 256);
    } else {
        if (lc == 64) NBH_GWQ(4, 64); else NBH_GWQ(4, 128);
    }
#undef NBH_GWQ
#undef NBH_GWE
#undef NBH_GWC
#undef NBH_GWL
#undef NBH_GWF
}
