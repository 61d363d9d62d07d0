/*
 * Tree construction on gfx950: root-size update, Morton keys, LSD radix sort,
 * Karras binary radix tree (as the Cartesian tree of the adjacent prefixes),
 * octree emission in pre-order and level-by-level
 * summarisation (centre of mass, critical radius, quadrupole).
 *
 * The octree is exactly the reference's (nbody/src/nbody_tree.c): root at the
 * origin with side st->tree.rsize, midpoints at exact dyadic positions, and a
 * cell wherever two or more massive bodies share a prefix -- including the
 * single-child chain cells that serial insertion creates (nbLoadBody,
 * :242-282).  Cells are the prefixes (level l, first body a) of the sorted
 * 189-bit keys; with delta_a the common prefix of sorted bodies a and a+1,
 * the cells whose first body is a are levels L(delta_{a-1})+1 .. L(delta_a),
 * L = floor(bits/3).  Pre-order (More/Next threading, :127-150) lists, for
 * each sorted body a, its new cells by level, then the body itself; so
 * pre(body a) = a + #cells with first body <= a and Next(cell [lo,hi]) =
 * pre(body hi) + 1.
 *
 * Several independent simulations can share one state: the simulation id
 * sits above the Morton bits of the sort key, prefixes across simulations
 * are -1, so every phase is one launch for all simulations, the binary radix
 * tree is a forest joined by prefix -1 nodes (which own no cell) and each
 * simulation's pre-order table is one contiguous segment.
 */
#include <cstring>

#include "nbhip_device.h"

#include <cfloat>

/* ---------------------------------------------------------------------- */
/* bounding box (+ optional opening half-kick and drift)                  */
/* ---------------------------------------------------------------------- */
/* advancePosVel (nbody/src/nbody_plain.c:268-284) fused with expandBox's
 * max-|x| scan (nbody/src/nbody_tree.c:156-176), per simulation.  Each
 * block owns a contiguous chunk of NBH_BBOX_CHUNK bodies (coalesced, 256
 * apart per thread); a thread flushes its running maximum only when its
 * simulation changes, and waves then the block combine theirs, so a launch
 * issues about one atomic per block and simulation instead of one per wave.
 * A NaN coordinate counts as infinite so the root-size step reports it. */
#define NBH_BBOX_CHUNK (256 * 16)

__device__ __forceinline__ void nbh_max_bits(SimTab* sims, int sim, double v)
{
    atomicMax(&sims[sim].xyzmax_bits, (unsigned long long) __double_as_longlong(v));
}

__global__ __launch_bounds__(256) void k_drift_bbox(int n, int drift, double dt, double dtHalf,
                                                    double* __restrict__ px, double* __restrict__ py,
                                                    double* __restrict__ pz, double* __restrict__ vx,
                                                    double* __restrict__ vy, double* __restrict__ vz,
                                                    const double* __restrict__ ax,
                                                    const double* __restrict__ ay,
                                                    const double* __restrict__ az,
                                                    const double* __restrict__ m,
                                                    double4* __restrict__ pos4,
                                                    const int* __restrict__ simid,
                                                    SimTab* __restrict__ sims,
                                                    const DevScalars* __restrict__ scal)
{
    __shared__ int s_sim[4];
    __shared__ double s_mx[4];
    if (scal->halt) return;               /* a fatal status stopped the run */
    const int lo = blockIdx.x * NBH_BBOX_CHUNK;
    const int hi = min(n, lo + NBH_BBOX_CHUNK);
    int cur = -1;
    double mx = 0.0;
    for (int i = lo + (int) threadIdx.x; i < hi; i += 256) {
        double x = px[i], y = py[i], z = pz[i];
        if (drift) {
            double ux = vx[i] + dtHalf * (ax[i] + 0.0);
            double uy = vy[i] + dtHalf * (ay[i] + 0.0);
            double uz = vz[i] + dtHalf * (az[i] + 0.0);
            x = x + dt * ux;
            y = y + dt * uy;
            z = z + dt * uz;
            vx[i] = ux; vy[i] = uy; vz[i] = uz;
            px[i] = x; py[i] = y; pz[i] = z;
        }
        pos4[i] = make_double4(x, y, z, m[i]);     /* gathered by the tree build */
        const int sim = simid[i];
        if (sim != cur) {
            if (cur >= 0) nbh_max_bits(sims, cur, mx);
            cur = sim;
            mx = 0.0;
        }
        double a = fmax(fabs(x), fmax(fabs(y), fabs(z)));
        if (x != x || y != y || z != z) a = INFINITY;
        mx = fmax(mx, a);
    }
    /* wave: one simulation among the lanes that hold one -> combine */
    const bool has = cur >= 0;
    const int w = threadIdx.x >> 6;
    const int sim_l0 = __builtin_amdgcn_readfirstlane(cur);
    const bool uniform = __ballot(has && cur != sim_l0) == 0ull;
    if (uniform) {
        const double m = nbh_wave_max(has ? mx : 0.0);
        if (__lane_id() == 0) {
            s_sim[w] = has ? cur : -1;
            s_mx[w] = m;
        }
    } else {
        if (has) nbh_max_bits(sims, cur, mx);
        if (__lane_id() == 0) s_sim[w] = -1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int sim = -1;
        double m = 0.0;
        for (int k = 0; k < 4; ++k) {
            if (s_sim[k] < 0) continue;
            if (s_sim[k] != sim) {
                if (sim >= 0) nbh_max_bits(sims, sim, m);
                sim = s_sim[k];
                m = 0.0;
            }
            m = fmax(m, s_mx[k]);
        }
        if (sim >= 0) nbh_max_bits(sims, sim, m);
    }
}

/* Root doubling rule (expandBox, nbody/src/nbody_tree.c:173-175) per
 * simulation, and per-build resets. */
__global__ __launch_bounds__(256) void k_rsize(int nsim, SimTab* __restrict__ sims,
                                               DevScalars* __restrict__ scal)
{
    /* Runs on a halted state too: the per-build counters below must start
     * every build from zero (a halted step skips the drift, so the maximum
     * read here is 0 and the root size does not change). */
    for (int k = threadIdx.x; k < nsim; k += 256) {
        SimTab& t = sims[k];
        const double xm = __longlong_as_double((long long) t.xyzmax_bits);
        double r = t.rsize;
        if (!(xm <= DBL_MAX)) {
            t.flags |= NBHIP_ERROR;
        } else {
            while (r < 2.0 * xm) r *= 2.0;
        }
        t.rsize = r;
        t.key_scale = 9223372036854775808.0 / r;
        t.xyzmax_bits = 0ull;
        t.first = -1;
        t.root = -1;
        t.end = -1;
        t.maxdepth = 0;
    }
    if (threadIdx.x == 0) {
        scal->nmassive = 0;
        scal->ncell = 0;
        scal->nnode = 0;
        scal->maxdepth = 0;
        scal->counters[0] = scal->counters[1] = scal->counters[2] = scal->counters[3] = 0ull;
        scal->counters[4] = 0ull;
    }
    if (threadIdx.x < 64) {
        scal->level_count[threadIdx.x] = 0;
        scal->level_fill[threadIdx.x] = 0;
    }
}

/* ---------------------------------------------------------------------- */
/* keys                                                                   */
/* ---------------------------------------------------------------------- */
/* Sort word: simulation id above the top `top` Morton levels.  Massless
 * bodies are not inserted (nbMakeTree, nbody/src/nbody_tree.c:438-442):
 * their key sorts after every massive key. */
__global__ __launch_bounds__(256) void k_keys(int n, const double* __restrict__ px,
                                              const double* __restrict__ py,
                                              const double* __restrict__ pz,
                                              const double* __restrict__ m,
                                              const int* __restrict__ simid,
                                              const SimTab* __restrict__ sims, int sim_shift,
                                              int top, unsigned long long* __restrict__ key,
                                              unsigned* __restrict__ perm)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    unsigned long long k = ~0ull;
    if (m[i] != 0.0) {
        const int sim = simid[i];
        const double ks = sims[sim].key_scale;
        const int sh = 63 - top;
        unsigned long long ux = nbh_fixed(px[i], ks);
        unsigned long long uy = nbh_fixed(py[i], ks);
        unsigned long long uz = nbh_fixed(pz[i], ks);
        k = nbh_morton21(ux >> sh, uy >> sh, uz >> sh);
        if (sim_shift < 64) k |= (unsigned long long) sim << sim_shift;
    }
    key[i] = k;
    perm[i] = (unsigned) i;
}

/* Massive keys are < 2^63, massless keys all-ones: after the sort the
 * massive bodies are the prefix [0, M). */
__global__ __launch_bounds__(256) void k_count_massive(int n,
                                                       const unsigned long long* __restrict__ key,
                                                       DevScalars* __restrict__ scal)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const bool massive = key[i] != ~0ull;
    if (i == 0 && !massive) scal->nmassive = 0;
    if (massive && (i + 1 == n || key[i + 1] == ~0ull)) scal->nmassive = i + 1;
}

/* ---------------------------------------------------------------------- */
/* LSD radix sort: 8-bit digits, 256 threads x 16 keys per block, stable   */
/* wave64 multisplit ranking.                                              */
/* ---------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_radix_hist(int n, int shift,
                                                    const unsigned long long* __restrict__ key,
                                                    unsigned* __restrict__ hist, int nb)
{
    __shared__ unsigned cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * NBH_SORT_TILE;
    /* all of the thread's keys in flight at once, then the counts */
    unsigned dg[NBH_SORT_TILE / 256];
    #pragma unroll
    for (int r = 0; r < NBH_SORT_TILE / 256; ++r) {
        const int idx = base + r * 256 + threadIdx.x;
        dg[r] = idx < n ? (unsigned) (key[idx] >> shift) & 255u : 256u;
    }
    #pragma unroll
    for (int r = 0; r < NBH_SORT_TILE / 256; ++r)
        if (dg[r] < 256u) atomicAdd(&cnt[dg[r]], 1u);
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];
}

/* Exclusive scan of each digit's row of the digit-major histogram (block d
 * owns digit d); the row total goes to dtot[d]. */
__global__ __launch_bounds__(256) void k_radix_scan(unsigned* __restrict__ hist,
                                                    unsigned* __restrict__ dtot, int nb)
{
    __shared__ unsigned wsum[4];
    unsigned* row = hist + (size_t) blockIdx.x * nb;
    const int t = threadIdx.x;
    const int chunk = (nb + 255) / 256;
    const int b0 = t * chunk;
    unsigned s = 0;
    for (int k = 0; k < chunk; ++k)
        if (b0 + k < nb) s += row[b0 + k];
    unsigned v = s;
    for (int off = 1; off < 64; off <<= 1) {
        unsigned u = __shfl_up(v, off, 64);
        if ((int) __lane_id() >= off) v += u;
    }
    if (__lane_id() == 63) wsum[t >> 6] = v;
    __syncthreads();
    unsigned before = 0;
    for (int w = 0; w < (t >> 6); ++w) before += wsum[w];
    unsigned excl = v - s + before;
    for (int k = 0; k < chunk; ++k) {
        if (b0 + k < nb) {
            unsigned h = row[b0 + k];
            row[b0 + k] = excl;
            excl += h;
        }
    }
    if (t == 255) dtot[blockIdx.x] = v + before;
}

/* Stable scatter of one pass.  Wave w of the block owns the tile's keys
 * [w * T/4, (w + 1) * T/4) as T/256 rows of 64; it ranks each row with a
 * wave64 multisplit (one ballot per digit bit) on top of its own per-digit
 * counts in LDS (the row's leader lane per digit reads and advances them,
 * so no barrier is needed between rows), then the four waves' counts are
 * scanned per digit.  Ranks therefore follow key order within the tile.
 * The tile is then laid out in digit order in LDS and written run by run
 * (each digit's run at the digit's base plus the block's offset, from the
 * histogram scan), so stores to the output are contiguous per run. */
__global__ __launch_bounds__(256) void k_radix_scatter(int n, int shift,
                                                       const unsigned long long* __restrict__ kin,
                                                       const unsigned* __restrict__ vin,
                                                       unsigned long long* __restrict__ kout,
                                                       unsigned* __restrict__ vout,
                                                       const unsigned* __restrict__ hist,
                                                       const unsigned* __restrict__ dtot, int nb)
{
    constexpr int R = NBH_SORT_TILE / 256;           /* rows per wave */
    __shared__ unsigned goff[256];
    __shared__ unsigned whist[4][256];
    __shared__ unsigned tbase[256];
    __shared__ unsigned wsum[4];
    __shared__ unsigned long long skey[NBH_SORT_TILE];
    __shared__ unsigned sval[NBH_SORT_TILE];
    const int t = threadIdx.x;
    const int w = t >> 6;
    const int lane = t & 63;
    /* digit bases: exclusive scan of the 256 digit totals */
    {
        const unsigned s0 = dtot[t];
        unsigned v = s0;
        for (int off = 1; off < 64; off <<= 1) {
            unsigned u = __shfl_up(v, off, 64);
            if (lane >= off) v += u;
        }
        if (lane == 63) wsum[w] = v;
        whist[0][t] = whist[1][t] = whist[2][t] = whist[3][t] = 0;
        __syncthreads();
        unsigned before = 0;
        for (int ww = 0; ww < w; ++ww) before += wsum[ww];
        goff[t] = v - s0 + before + hist[t * nb + blockIdx.x];
    }
    const unsigned long long lt = nbh_lanemask_lt();
    const int base = blockIdx.x * NBH_SORT_TILE + w * (NBH_SORT_TILE / 4);
    unsigned long long kr[R];
    unsigned vr[R];
    unsigned rk[R];
    #pragma unroll
    for (int r = 0; r < R; ++r) {
        const int idx = base + r * 64 + lane;
        kr[r] = idx < n ? kin[idx] : 0ull;
        vr[r] = idx < n ? vin[idx] : 0u;
    }
    unsigned* wh = whist[w];
    #pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool valid = base + r * 64 + lane < n;
        const unsigned d = (unsigned) (kr[r] >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
        #pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long bb = __ballot(valid && ((d >> b) & 1u));
            peers &= ((d >> b) & 1u) ? bb : ~bb;
        }
        const int leader = valid ? __builtin_ctzll(peers) : lane;
        unsigned before = 0;
        if (valid && lane == leader) {
            before = wh[d];
            wh[d] = before + (unsigned) __popcll(peers);
        }
        before = (unsigned) __shfl((int) before, leader, 64);
        rk[r] = before + (unsigned) __popcll(peers & lt);
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    /* digit d = t: the waves' offsets within the tile's run of digit d, and
     * the run's start in the tile (an exclusive scan over digits) */
    unsigned acc = 0;
    #pragma unroll
    for (int ww = 0; ww < 4; ++ww) {
        const unsigned c = whist[ww][t];
        whist[ww][t] = acc;
        acc += c;
    }
    {
        unsigned v = acc;
        for (int off = 1; off < 64; off <<= 1) {
            unsigned u = __shfl_up(v, off, 64);
            if (lane >= off) v += u;
        }
        if (lane == 63) wsum[w] = v;
        __syncthreads();
        unsigned before = 0;
        for (int ww = 0; ww < w; ++ww) before += wsum[ww];
        tbase[t] = v - acc + before;
    }
    __syncthreads();
    /* the tile in digit order in LDS, then written out run by run: lanes
     * that follow each other write neighbouring addresses of one run */
    #pragma unroll
    for (int r = 0; r < R; ++r) {
        if (base + r * 64 + lane >= n) continue;
        const unsigned d = (unsigned) (kr[r] >> shift) & 255u;
        const unsigned lp = tbase[d] + wh[d] + rk[r];
        skey[lp] = kr[r];
        sval[lp] = vr[r];
    }
    __syncthreads();
    const int cnt = min(NBH_SORT_TILE, n - blockIdx.x * NBH_SORT_TILE);
    for (int i = t; i < cnt; i += 256) {
        const unsigned long long k = skey[i];
        const unsigned d = (unsigned) (k >> shift) & 255u;
        const unsigned pos = goff[d] + ((unsigned) i - tbase[d]);
        kout[pos] = k;
        vout[pos] = sval[i];
    }
}

/* ---------------------------------------------------------------------- */
/* ties in the sort word: order by the remaining levels                    */
/* ---------------------------------------------------------------------- */
/* Morton word of the 21 levels that start at level a (1-based); levels past
 * 63 read as zero. */
__device__ __forceinline__ unsigned long long nbh_word_at(unsigned long long ux,
                                                          unsigned long long uy,
                                                          unsigned long long uz, int a)
{
    const int s = a - 1;
    return nbh_morton21(((ux << 1) << s) >> 43, ((uy << 1) << s) >> 43, ((uz << 1) << s) >> 43);
}

__device__ __forceinline__ void nbh_deep_words(const double* px, const double* py,
                                               const double* pz, double ks, unsigned j, int top,
                                               unsigned long long* w)
{
    unsigned long long ux = nbh_fixed(px[j], ks), uy = nbh_fixed(py[j], ks),
                       uz = nbh_fixed(pz[j], ks);
    w[0] = nbh_word_at(ux, uy, uz, top + 1);
    w[1] = nbh_word_at(ux, uy, uz, top + 22);
    w[2] = top + 43 <= 63 ? nbh_word_at(ux, uy, uz, top + 43) : 0ull;
}

__device__ __forceinline__ bool nbh_words_gt(const unsigned long long* a,
                                             const unsigned long long* b)
{
    if (a[0] != b[0]) return a[0] > b[0];
    if (a[1] != b[1]) return a[1] > b[1];
    return a[2] > b[2];
}

__global__ __launch_bounds__(256) void k_tiefix_rank(const unsigned long long* __restrict__ key,
                                                     const unsigned* __restrict__ perm,
                                                     unsigned* __restrict__ perm_out,
                                                     const double* __restrict__ px,
                                                     const double* __restrict__ py,
                                                     const double* __restrict__ pz,
                                                     const int* __restrict__ simid,
                                                     const SimTab* __restrict__ sims, int top,
                                                     const DevScalars* __restrict__ scal)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int M = scal->nmassive;
    if (i >= M) return;
    const unsigned long long k = key[i];
    const bool tie = (i > 0 && key[i - 1] == k) || (i + 1 < M && key[i + 1] == k);
    if (!tie) return;
    int a = i;
    while (a > 0 && key[a - 1] == k) --a;
    int e = i + 1;
    while (e < M && key[e] == k) ++e;
    const double ks = sims[simid[perm[i]]].key_scale;
    unsigned long long wi[3];
    nbh_deep_words(px, py, pz, ks, perm[i], top, wi);
    /* rank in the run: members with smaller deeper words, and members with
     * equal words that come earlier (the stable order of the sort) */
    int rank = 0;
    for (int j = a; j < e; ++j) {
        if (j == i) continue;
        unsigned long long wj[3];
        nbh_deep_words(px, py, pz, ks, perm[j], top, wj);
        if (nbh_words_gt(wi, wj) || (j < i && !nbh_words_gt(wj, wi))) ++rank;
    }
    perm_out[a + rank] = perm[i];
}

__global__ __launch_bounds__(256) void k_tiefix_copy(const unsigned long long* __restrict__ key,
                                                     unsigned* __restrict__ perm,
                                                     const unsigned* __restrict__ perm_out,
                                                     const DevScalars* __restrict__ scal)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int M = scal->nmassive;
    if (i >= M) return;
    const unsigned long long k = key[i];
    if ((i > 0 && key[i - 1] == k) || (i + 1 < M && key[i + 1] == k)) perm[i] = perm_out[i];
}

/* ---------------------------------------------------------------------- */
/* sorted gather, adjacent prefixes, cells-per-first-body                  */
/* ---------------------------------------------------------------------- */
__global__ __launch_bounds__(256) void k_gather(int n, const unsigned* __restrict__ perm,
                                                const double4* __restrict__ pos4,
                                                const unsigned long long* __restrict__ key,
                                                int sim_shift,
                                                const int* __restrict__ simid,
                                                const SimTab* __restrict__ sims,
                                                double4* __restrict__ spos,
                                                int* __restrict__ simid_s,
                                                unsigned long long* __restrict__ su,
                                                const DevScalars* __restrict__ scal)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const unsigned j = perm[s];
    /* one 32-byte record per body; the simulation of a massive body is the
     * top of its sorted key (massless bodies carry an all-ones key) */
    const double4 p = pos4[j];
    const double x = p.x, y = p.y, z = p.z;
    const bool massive = s < scal->nmassive;
    const int sim = (sim_shift >= 64) ? 0 : (massive ? (int) (key[s] >> sim_shift) : simid[j]);
    spos[s] = p;
    simid_s[s] = sim;
    if (massive) {
        const double ks = sims[sim].key_scale;
        su[3 * (size_t) s + 0] = nbh_fixed(x, ks);
        su[3 * (size_t) s + 1] = nbh_fixed(y, ks);
        su[3 * (size_t) s + 2] = nbh_fixed(z, ks);
    }
}

/* Common prefix, in bits, of the full keys (simulation id in the top sb
 * bits, then the 189-bit Morton key) of sorted massive bodies i and j; -1
 * outside [0, M); equal keys ordered by index (Karras 2012, section 4). */
__device__ __forceinline__ int nbh_kdelta(const unsigned long long* __restrict__ su,
                                          const int* __restrict__ simid_s, int n, int M, int sb,
                                          int i, int j)
{
    if (j < 0 || j >= M) return -1;
    const int si = simid_s[i], sj = simid_s[j];
    if (si != sj) return __clz((unsigned) (si ^ sj)) - (32 - sb);
    const unsigned long long* a = su + 3 * (size_t) i;
    const unsigned long long* b = su + 3 * (size_t) j;
    int d = nbh_prefix_bits(a[0], a[1], a[2], b[0], b[1], b[2]);
    if (d == 189) d = 189 + __clz((unsigned) (i ^ j));
    return sb + d;
}

__device__ __forceinline__ bool nbh_first_of_sim(const int* __restrict__ simid_s, int a)
{
    return a == 0 || simid_s[a - 1] != simid_s[a];
}

__global__ __launch_bounds__(256) void k_delta_count(int n, const unsigned long long* __restrict__ su,
                                                     const int* __restrict__ simid_s,
                                                     int* __restrict__ dadj,
                                                     unsigned* __restrict__ ccount,
                                                     SimTab* __restrict__ sims,
                                                     const DevScalars* __restrict__ scal)
{
    __shared__ int bsim, bmax, mixed;
    __shared__ int s_min16[16];
    if (threadIdx.x == 0) {
        bsim = -1;
        bmax = -1;
        mixed = 0;
    }
    __syncthreads();
    const int a = blockIdx.x * 256 + threadIdx.x;
    const int M = scal->nmassive;
    const int sb = scal->sim_bits;
    int depth = -1, sim = -1;
    int hv = INT_MAX;                  /* this prefix in the min-hierarchy (none: +inf) */
    if (a < M) {
        const int d = nbh_kdelta(su, simid_s, n, M, sb, a, a + 1);
        dadj[a] = d;
        hv = d;
        int c;
        if (nbh_first_of_sim(simid_s, a)) {
            c = max(nbh_levels(d, sb), 0) + 1;
        } else {
            const int dp = nbh_kdelta(su, simid_s, n, M, sb, a - 1, a);
            c = max(0, nbh_levels(d, sb) - nbh_levels(dp, sb));
        }
        ccount[a] = (unsigned) c;
        depth = max(nbh_levels(d, sb), 0);
        sim = simid_s[a];
        if (threadIdx.x == 0) bsim = sim;
    }
    /* the min-hierarchy's levels 1 and 2: minima of 16 and of 256 prefixes */
    for (int o = 8; o > 0; o >>= 1) hv = min(hv, __shfl_xor(hv, o, 16));
    const NbhDHier h = nbh_dhier(n, M);
    if ((threadIdx.x & 15) == 0) {
        if (h.top >= 1 && (a >> 4) < h.size[1]) dadj[h.off[1] + (a >> 4)] = hv;
        s_min16[threadIdx.x >> 4] = hv;
    }
    __syncthreads();
    if (threadIdx.x == 0 && h.top >= 2 && (int) blockIdx.x < h.size[2]) {
        int bm = INT_MAX;
        for (int k = 0; k < 16; ++k) bm = min(bm, s_min16[k]);
        dadj[h.off[2] + blockIdx.x] = bm;
    }
    if (sim >= 0 && sim != bsim) mixed = 1;
    /* the workgroup's deepest level: a wave maximum, then one LDS atomic per
     * wave (not one per lane on a single address) */
    const int wmax = nbh_wave_max_i(sim >= 0 ? depth : -1);
    if (__lane_id() == 0 && wmax >= 0) atomicMax(&bmax, wmax);
    __syncthreads();
    /* the maximum only grows, so a value already at least as deep (even a
     * stale read of it) makes the atomic unnecessary: few of the workgroups
     * that share a simulation reach it */
    if (sim >= 0 && mixed && depth > sims[sim].maxdepth) atomicMax(&sims[sim].maxdepth, depth);
    if (threadIdx.x == 0 && !mixed && bsim >= 0 && bmax > sims[bsim].maxdepth)
        atomicMax(&sims[bsim].maxdepth, bmax);
}

/* ---------------------------------------------------------------------- */
/* inclusive scan of cells-per-first-body (length nmassive)               */
/* ---------------------------------------------------------------------- */
__device__ __forceinline__ unsigned nbh_block1024_excl(unsigned s, unsigned* wsum,
                                                       unsigned* total)
{
    const int t = threadIdx.x;
    unsigned v = s;
    for (int off = 1; off < 64; off <<= 1) {
        unsigned u = __shfl_up(v, off, 64);
        if ((int) __lane_id() >= off) v += u;
    }
    if (__lane_id() == 63) wsum[t >> 6] = v;
    __syncthreads();
    if (t < 16) {
        unsigned w = wsum[t];
        for (int off = 1; off < 16; off <<= 1) {
            unsigned u = __shfl_up(w, off, 16);
            if (t >= off) w += u;
        }
        wsum[t] = w;
    }
    __syncthreads();
    *total = wsum[15];
    return v - s + ((t >> 6) ? wsum[(t >> 6) - 1] : 0u);
}

__global__ __launch_bounds__(1024) void k_scan_tiles(unsigned* __restrict__ data,
                                                     unsigned* __restrict__ part,
                                                     const DevScalars* __restrict__ scal)
{
    __shared__ unsigned wsum[16];
    const int M = scal->nmassive;
    const int base = blockIdx.x * NBH_SCAN_TILE + threadIdx.x * 4;
    if (blockIdx.x * NBH_SCAN_TILE >= M) return;
    unsigned v[4];
    #pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = (base + k < M) ? data[base + k] : 0u;
    unsigned s = v[0] + v[1] + v[2] + v[3];
    unsigned tot;
    unsigned ex = nbh_block1024_excl(s, wsum, &tot);
    #pragma unroll
    for (int k = 0; k < 4; ++k) {
        ex += v[k];
        if (base + k < M) data[base + k] = ex;
    }
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_scan_part(unsigned* __restrict__ part,
                                                    const DevScalars* __restrict__ scal)
{
    __shared__ unsigned wsum[16];
    const int M = scal->nmassive;
    const int nb = (M + NBH_SCAN_TILE - 1) / NBH_SCAN_TILE;
    const int base = threadIdx.x * 4;
    unsigned v[4];
    #pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = (base + k < nb) ? part[base + k] : 0u;
    unsigned s = v[0] + v[1] + v[2] + v[3];
    unsigned tot;
    unsigned ex = nbh_block1024_excl(s, wsum, &tot);
    #pragma unroll
    for (int k = 0; k < 4; ++k) {
        ex += v[k];
        if (base + k < nb) part[base + k] = ex;
    }
}

__global__ __launch_bounds__(1024) void k_scan_add(unsigned* __restrict__ data,
                                                   const unsigned* __restrict__ part,
                                                   const DevScalars* __restrict__ scal)
{
    const int M = scal->nmassive;
    if (blockIdx.x == 0) return;
    const int base = blockIdx.x * NBH_SCAN_TILE + threadIdx.x * 4;
    const unsigned add = part[blockIdx.x - 1];
    #pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + k < M) data[base + k] += add;
}

/* Exclusive scan of the children per node over the node table (length
 * scal->nnode): each cell's child block in the group walk table.  Tiles of
 * NBH_SCAN_TILE entries, then the tile totals (one workgroup, 4096 at a time
 * with a running carry), then the tile offsets added back.  A node table
 * that overflowed its capacity has no child counts: nothing is scanned. */
__device__ __forceinline__ int nbh_nscan_len(const DevScalars* scal, int node_cap)
{
    if (scal->status & NBHIP_CELL_OVERFLOW_ERROR) return 0;
    return min(scal->nnode, node_cap);
}

__global__ __launch_bounds__(1024) void k_nscan_tiles(const unsigned* __restrict__ in,
                                                      unsigned* __restrict__ out,
                                                      unsigned* __restrict__ part,
                                                      int node_cap,
                                                      const DevScalars* __restrict__ scal)
{
    __shared__ unsigned wsum[16];
    const int nn = nbh_nscan_len(scal, node_cap);
    if (blockIdx.x * NBH_SCAN_TILE >= nn) return;
    const int base = blockIdx.x * NBH_SCAN_TILE + threadIdx.x * 4;
    unsigned v[4];
    #pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = (base + k < nn) ? (unsigned) __popc(in[base + k]) : 0u;
    const unsigned sum = v[0] + v[1] + v[2] + v[3];
    unsigned tot;
    unsigned ex = nbh_block1024_excl(sum, wsum, &tot);
    #pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < nn) out[base + k] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_nscan_part(unsigned* __restrict__ part, int node_cap,
                                                     const DevScalars* __restrict__ scal)
{
    __shared__ unsigned wsum[16];
    const int nb = (nbh_nscan_len(scal, node_cap) + NBH_SCAN_TILE - 1) / NBH_SCAN_TILE;
    unsigned carry = 0;
    for (int c0 = 0; c0 < nb; c0 += NBH_SCAN_TILE) {
        const int base = c0 + threadIdx.x * 4;
        unsigned v[4];
        #pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = (base + k < nb) ? part[base + k] : 0u;
        const unsigned sum = v[0] + v[1] + v[2] + v[3];
        unsigned tot;
        unsigned ex = carry + nbh_block1024_excl(sum, wsum, &tot);
        #pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (base + k < nb) part[base + k] = ex;      /* exclusive tile offsets */
            ex += v[k];
        }
        carry += tot;
        __syncthreads();          /* wsum is reused by the next chunk */
    }
}

__global__ __launch_bounds__(1024) void k_nscan_add(unsigned* __restrict__ out,
                                                    const unsigned* __restrict__ part,
                                                    int node_cap,
                                                    const DevScalars* __restrict__ scal)
{
    const int nn = nbh_nscan_len(scal, node_cap);
    if (blockIdx.x == 0 || blockIdx.x * NBH_SCAN_TILE >= nn) return;
    const int base = blockIdx.x * NBH_SCAN_TILE + threadIdx.x * 4;
    const unsigned add = part[blockIdx.x];
    #pragma unroll
    for (int k = 0; k < 4; ++k)
        if (base + k < nn) out[base + k] += add;
}

/* ---------------------------------------------------------------------- */
/* node counts, per-simulation segments                                    */
/* ---------------------------------------------------------------------- */

/* Each simulation's first body (which owns its root cell) and root; thread 0 also finalises the cell
 * and node counts (cells = the inclusive scan's last entry) and flags a node
 * table overflow. */
__global__ __launch_bounds__(256) void k_sim_roots(const int* __restrict__ simid_s,
                                                   const unsigned* __restrict__ cincl,
                                                   int node_cap, SimTab* __restrict__ sims,
                                                   unsigned* __restrict__ nkid,
                                                   DevScalars* __restrict__ scal)
{
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a == 0) {
        const int M = scal->nmassive;
        const int C = M > 0 ? (int) cincl[M - 1] : 0;
        const int nn = M + C;
        scal->ncell = C;
        scal->nnode = nn;
        if (nn > node_cap) scal->status |= NBHIP_CELL_OVERFLOW_ERROR;
    }
    if (a >= scal->nmassive) return;
    {
        /* child counters of body a's new cells and its leaf (pre-order
         * a + cincl[a-1] .. a + cincl[a]), counted by the emission */
        const int lo = a + (a > 0 ? (int) cincl[a - 1] : 0);
        const int hi = a + (int) cincl[a];
        for (int i = lo; i <= hi && i < node_cap; ++i) nkid[i] = 0u;
    }
    if (!nbh_first_of_sim(simid_s, a)) return;
    const int k = simid_s[a];
    sims[k].first = a;
    sims[k].root = a + (a > 0 ? (int) cincl[a - 1] : 0);
}

__global__ __launch_bounds__(256) void k_sim_ends(int nsim, SimTab* __restrict__ sims,
                                                  DevScalars* __restrict__ scal)
{
    for (int k = threadIdx.x; k < nsim; k += 256) {
        if (sims[k].root < 0) continue;
        int end = scal->nnode;
        for (int j = k + 1; j < nsim; ++j)
            if (sims[j].root >= 0) {
                end = sims[j].root;
                break;
            }
        sims[k].end = end;
        atomicMax(&scal->maxdepth, sims[k].maxdepth);
    }
}

__device__ __forceinline__ double nbh_rcrit2(const nbhip_ctx& c, double mx, double my, double mz,
                                             double rsize, double cx, double cy, double cz,
                                             double psize)
{
    /* findRCrit (nbody/src/nbody_tree.c:308-343) */
    double rc;
    if (c.theta == 0.0) {
        rc = 2.0 * rsize;
        return rc * rc;
    }
    if (c.criterion == NBHIP_TREECODE) {
        double dx = cx - mx, dy = cy - my, dz = cz - mz;
        rc = psize / c.theta + sqrt(dx * dx + dy * dy + dz * dz);
        return rc * rc;
    }
    if (c.criterion == NBHIP_SW93) {
        double d0 = cx - (mx - 0.5 * psize);
        double e0 = fmax(d0, psize - d0);
        double bmax2 = e0 * e0;
        double d1 = cy - (my - 0.5 * psize);
        double e1 = fmax(d1, psize - d1);
        bmax2 += e1 * e1;
        double d2 = cz - (mz - 0.5 * psize);
        double e2 = fmax(d2, psize - d2);
        bmax2 += e2 * e2;
        return bmax2 / (c.theta * c.theta);
    }
    rc = psize / c.theta;
    return rc * rc;
}

/* ---------------------------------------------------------------------- */
/* Karras binary radix tree (a forest across simulations)                  */
/* ---------------------------------------------------------------------- */
/* Levels 3 and up of the adjacent prefixes' min-hierarchy (levels 1 and 2
 * come from k_delta_count), one workgroup, level by level. */
__global__ __launch_bounds__(1024) void k_dhier_top(int n, int* __restrict__ dadj,
                                                    const DevScalars* __restrict__ scal)
{
    __shared__ NbhDHier h;
    if (threadIdx.x == 0) h = nbh_dhier(n, scal->nmassive);
    __syncthreads();
    for (int L = 3; L <= h.top; ++L) {
        for (int i = threadIdx.x; i < h.size[L]; i += 1024) {
            int m = INT_MAX;
            for (int t = 0; t < 16; ++t) {
                const int c = 16 * i + t;
                if (c < h.size[L - 1]) m = min(m, dadj[h.off[L - 1] + c]);
            }
            dadj[h.off[L] + i] = m;
        }
        __syncthreads();
    }
}

/* Entries base .. base+15 of hierarchy level L below v (bit t: entry
 * base + t), entries past the level's end excluded; base is a multiple of
 * 16, so the chunk is four aligned 16-byte loads. */
__device__ __forceinline__ unsigned nbh_dhier_below(const int* __restrict__ dadj, const NbhDHier& h,
                                                    int L, int base, int v)
{
    const int4* q = reinterpret_cast<const int4*>(dadj + h.off[L] + base);
    unsigned m = 0;
    #pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int4 e = q[k];
        m |= (unsigned) (e.x < v) << (4 * k) | (unsigned) (e.y < v) << (4 * k + 1) |
             (unsigned) (e.z < v) << (4 * k + 2) | (unsigned) (e.w < v) << (4 * k + 3);
    }
    const int lim = h.size[L] - base;
    return lim >= 16 ? m : (m & ((1u << max(lim, 0)) - 1u));
}

/* Nearest adjacent prefix smaller than v = dadj[i]: to the right (the
 * first p > i; one exists, dadj[M-1] = -1) or to the left (the last p < i,
 * -1 if none).  Up the hierarchy through the siblings of the chunk passed,
 * then down into the first (last) chunk holding one: at most twice its
 * height in rounds of one 64-byte chunk.  -2 on an inconsistent hierarchy. */
__device__ __forceinline__ int nbh_nearest_smaller(const int* __restrict__ dadj, const NbhDHier& h,
                                                   int i, int v, bool right)
{
    int c = i, L = 0;
    for (;;) {
        const int base = c & ~15, t = c & 15;
        unsigned m = nbh_dhier_below(dadj, h, L, base, v);
        m &= right ? ~((2u << t) - 1u) : ((1u << t) - 1u);
        if (m) {
            c = base + (right ? __builtin_ctz(m) : 31 - __builtin_clz(m));
            break;
        }
        if (L >= h.top) return right ? -2 : -1;
        c >>= 4;
        ++L;
    }
    for (; L > 0; --L) {
        const unsigned m = nbh_dhier_below(dadj, h, L - 1, c << 4, v);
        if (!m) return -2;
        c = (c << 4) + (right ? __builtin_ctz(m) : 31 - __builtin_clz(m));
    }
    return c;
}

/* ---------------------------------------------------------------------- */
/* octree emission in pre-order                                            */
/* ---------------------------------------------------------------------- */
struct EmitView {
    const int* dadj;
    const unsigned* cincl;
    int sb;
};

__device__ __forceinline__ int nbh_first_level(const EmitView& v, int lo)
{
    return lo == 0 ? 0 : nbh_levels(v.dadj[lo - 1], v.sb) + 1;
}

__device__ __forceinline__ int nbh_excl(const EmitView& v, int lo)
{
    return lo == 0 ? 0 : (int) v.cincl[lo - 1];
}

/* Pre-order index of the level-P cell that contains sorted body i: its first
 * body a follows the nearest adjacent prefix left of i below level P (a
 * prefix reaches level P iff it is at least sb + 3P), and the cells whose
 * first body is a precede a's leaf level by level.  -1 on an inconsistent
 * hierarchy. */
__device__ __forceinline__ int nbh_first_body(const EmitView& v, const NbhDHier& h, int i, int P)
{
    return nbh_nearest_smaller(v.dadj, h, i, v.sb + 3 * P, false) + 1;
}
__device__ __forceinline__ int nbh_cell_pre(const EmitView& v, int a, int P)
{
    return a < 0 ? -1 : a + nbh_excl(v, a) + (P - nbh_first_level(v, a));
}
__device__ __forceinline__ int nbh_cell_at(const EmitView& v, const NbhDHier& h, int i, int P)
{
    return nbh_cell_pre(v, nbh_first_body(v, h, i, P), P);
}

/* Octant at level l (1..63) of a key: its level-l bit of each coordinate.
 * The children of an octree cell sit in distinct octants, so the octant is
 * a child's slot in its parent's child mask (nkid; a cell's child count is
 * the mask's population, its children's order in the child list the order
 * of their octants). */
__device__ __forceinline__ int nbh_octant(unsigned long long ux, unsigned long long uy,
                                          unsigned long long uz, int l)
{
    const int b = 63 - l;
    return (int) (((ux >> b) & 1ull) << 2 | ((uy >> b) & 1ull) << 1 | ((uz >> b) & 1ull));
}

__device__ __forceinline__ void nbh_emit_cell(int pre, int l, int next, int parent, int sim,
                                              double R, unsigned long long ux,
                                              unsigned long long uy, unsigned long long uz,
                                              double* __restrict__ nmid,
                                              int* __restrict__ nlevel, int* __restrict__ nparent,
                                              int* __restrict__ nsim, SimTab* __restrict__ sims,
                                              int* __restrict__ nnext)
{
    /* the cell's walk record is written whole by its summary; until then
     * its Next link lives in nnext */
    nnext[pre] = next;
    nlevel[pre] = l;
    nparent[pre] = parent;
    nsim[pre] = sim;
    double mx = 0.0, my = 0.0, mz = 0.0;
    if (l > 0) {
        const double h = ldexp(R, -(l + 1));
        const int sh = 63 - l;
        mx = (double) (2ull * (ux >> sh) + 1ull) * h - 0.5 * R;
        my = (double) (2ull * (uy >> sh) + 1ull) * h - 0.5 * R;
        mz = (double) (2ull * (uz >> sh) + 1ull) * h - 0.5 * R;
    }
    nmid[3 * pre + 0] = mx;
    nmid[3 * pre + 1] = my;
    nmid[3 * pre + 2] = mz;
    /* Structure-error rule of nbLoadBody (nbody/src/nbody_tree.c:256-264):
     * a body descends through a cell of side <= DBL_EPSILON with an occupied
     * slot, i.e. a cell exists at level l with rsize * 2^-(l-1) <= DBL_EPSILON. */
    if (l >= 1 && ldexp(R, -(l - 1)) <= DBL_EPSILON) atomicOr(&sims[sim].flags, NBHIP_TREE_STRUCTURE_ERROR);
    else if (l >= 63) atomicOr(&sims[sim].flags, NBHIP_MAX_DEPTH_ERROR);
}

/* The octree cells of split k of the binary radix tree, thread per split.
 * That tree is the minimum Cartesian tree of the adjacent prefixes dadj[0..
 * M-2] (a sorted key's common prefix with the next; the index tie-break of
 * equal keys puts every range's least prefix at one position; T. Karras,
 * HPG 2012, builds the same tree by binary searches over the keys): the
 * node split at k covers bodies L+1 .. R, L and R the nearest prefixes
 * smaller than dadj[k] to its left and right, and its parent is split at
 * whichever of the two is larger.  The node owns the octree levels of its
 * prefix below its parent's, each a cell [lo, hi] whose first body is lo;
 * the topmost one's parent is the level-Lp cell holding lo. */
__global__ __launch_bounds__(256) void k_emit_cells(int n, const int* __restrict__ dadj,
                                                    const unsigned* __restrict__ cincl,
                                                    const unsigned long long* __restrict__ su,
                                                    const int* __restrict__ simid_s,
                                                    double* __restrict__ nmid,
                                                    int* __restrict__ nlevel,
                                                    int* __restrict__ nparent,
                                                    int* __restrict__ nsim,
                                                    SimTab* __restrict__ sims,
                                                    unsigned* __restrict__ nkid,
                                                    int* __restrict__ nrank,
                                                    int* __restrict__ nnext,
                                                    DevScalars* __restrict__ scal)
{
    __shared__ NbhDHier h;             /* indexed by level: kept in LDS, not in scratch */
    const int k = blockIdx.x * 256 + threadIdx.x;
    const int M = scal->nmassive;
    if (threadIdx.x == 0) h = nbh_dhier(n, M);
    __syncthreads();
    if (k >= M - 1) return;
    if (scal->status & NBHIP_CELL_OVERFLOW_ERROR) return;
    const EmitView v = { dadj, cincl, scal->sim_bits };
    const int d = dadj[k];
    const int R = nbh_nearest_smaller(dadj, h, k, d, true);
    const int Lft = nbh_nearest_smaller(dadj, h, k, d, false);
    if (R < 0 || Lft < -1) {
        atomicOr(&scal->status, NBHIP_ERROR);
        return;
    }
    const int lo = Lft + 1, hi = R;
    const int Lk = nbh_levels(d, v.sb);
    const int Lp = nbh_levels(max(Lft >= 0 ? dadj[Lft] : -1, R < M - 1 ? dadj[R] : -1), v.sb);
    if (Lk <= Lp) return;
    const int base = lo + nbh_excl(v, lo) - nbh_first_level(v, lo);
    const int next = hi + 1 + (int) cincl[hi];
    const int sim = simid_s[lo];
    const double Rs = sims[sim].rsize;
    const unsigned long long ux = su[3 * (size_t) lo], uy = su[3 * (size_t) lo + 1],
                             uz = su[3 * (size_t) lo + 2];
    const int nn = scal->nnode;
    int top_parent = -1;
    if (Lp >= 0) {
        top_parent = nbh_cell_at(v, h, lo, Lp);
        if (top_parent < 0) {
            atomicOr(&scal->status, NBHIP_ERROR);
            return;
        }
    }
    for (int l = Lp + 1; l <= Lk; ++l) {
        const int pre = base + l;
        if (pre < 0 || pre >= nn) {
            atomicOr(&scal->status, NBHIP_ERROR);
            return;
        }
        const int par = (l == Lp + 1) ? top_parent : pre - 1;
        nbh_emit_cell(pre, l, next, par, sim, Rs, ux, uy, uz, nmid, nlevel, nparent, nsim, sims,
                      nnext);
        /* the cell's octant in its parent, and the parent's octant mask */
        if (par >= 0) {
            const int o = nbh_octant(ux, uy, uz, l);
            nrank[pre] = o;
            atomicOr(&nkid[par], 1u << o);
        }
    }
}

/* Body leaves; a simulation with a single massive body gets its root cell
 * here (it owns no binary node). */
__global__ __launch_bounds__(256) void k_emit_leaves(int n, const int* __restrict__ dadj,
                                                     const unsigned long long* __restrict__ su,
                                                     const unsigned* __restrict__ cincl,
                                                     const double4* __restrict__ spos,
                                                     const int* __restrict__ simid_s,
                                                     NodeW* __restrict__ node,
                                                     double* __restrict__ nmid,
                                                     int* __restrict__ nlevel,
                                                     int* __restrict__ nparent,
                                                     int* __restrict__ nsim,
                                                     SimTab* __restrict__ sims,
                                                     NodeF* __restrict__ nodef,
                                                     unsigned* __restrict__ nkid,
                                                     int* __restrict__ nrank,
                                                     int* __restrict__ nnext,
                                                     DevScalars* __restrict__ scal)
{
    __shared__ NbhDHier h;             /* indexed by level: kept in LDS, not in scratch */
    const int s = blockIdx.x * 256 + threadIdx.x;
    const int M = scal->nmassive;
    if (threadIdx.x == 0) h = nbh_dhier(n, M);
    __syncthreads();
    if (s >= M) return;
    if (scal->status & NBHIP_CELL_OVERFLOW_ERROR) return;
    const EmitView v = { dadj, cincl, scal->sim_bits };
    const int pre = s + (int) cincl[s];
    const int sim = simid_s[s];
    const bool singleton = nbh_first_of_sim(simid_s, s) && (s + 1 == M || simid_s[s + 1] != sim);
    int parent, P = 0, a = s;
    if (singleton) {
        parent = pre - 1;
        if (parent >= 0)
            nbh_emit_cell(parent, 0, pre + 1, -1, sim, sims[sim].rsize, 0ull, 0ull, 0ull, nmid,
                          nlevel, nparent, nsim, sims, nnext);
    } else {
        /* the deepest cell holding s: the deeper of its two prefixes' levels */
        P = max(nbh_levels(s > 0 ? dadj[s - 1] : -1, v.sb), nbh_levels(dadj[s], v.sb));
        a = nbh_first_body(v, h, s, P);
        parent = nbh_cell_pre(v, a, P);
    }
    if (pre < 0 || pre >= scal->nnode || parent < 0 || parent >= pre) {
        atomicOr(&scal->status, NBHIP_ERROR);
        return;
    }
    /* the leaf's slot in its parent: its octant one level down; in a
     * level-63 cell (equal keys, all children leaves) its place among them
     * (past 8: the overflow slot 8, which the child lists reject) */
    const int o = P < 63 ? nbh_octant(su[3 * (size_t) s], su[3 * (size_t) s + 1],
                                      su[3 * (size_t) s + 2], P + 1)
                         : min(s - a, 8);
    nrank[pre] = o;
    atomicOr(&nkid[parent], 1u << o);
    const double4 p = spos[s];
    NodeW b;
    b.x = p.x; b.y = p.y; b.z = p.z; b.m = p.w;
    b.rcrit2 = 0.0;
    b.next = pre + 1;
    b.body = s;
    node[pre] = b;
    nlevel[pre] = -1;
    nparent[pre] = parent;
    nsim[pre] = sim;
    if (nodef) {
        NodeF f;
        f.x = p.x;
        f.y = p.y;
        f.z = p.z;
        f.m = (float) p.w;
        f.rcrit2 = -1.0f;
        f.next = pre + 1;
        f.body = s;
        for (int k = 0; k < 6; ++k) f.q[k] = 0.0f;
        nodef[pre] = f;
    }
}

/* Cells grouped by level: each workgroup forms the exclusive scan of the
 * per-level counts itself (64 entries; workgroup 0 also stores it for the
 * summaries), then each cell takes a slot in its level's segment (order
 * within a level is irrelevant: cells of one level are independent). */

/* Each workgroup takes 4096 consecutive nodes at a time (grid-stride over
 * the nodes actually built): it counts their cells per level in LDS, claims
 * one range per non-empty level with a single global atomic, then places
 * them.  Claims per level stay few (a few hundred workgroups), so the 64
 * level counters see little contention. */
#define NBH_LS_CHUNK 4096
__global__ __launch_bounds__(256) void k_level_scatter(const int* __restrict__ nlevel,
                                                       int* __restrict__ level_list,
                                                       DevScalars* __restrict__ scal)
{
    __shared__ int lh[64];
    __shared__ int base[64];
    __shared__ int off[65];
    const int tid = threadIdx.x;
    __shared__ int consistent;
    if (tid == 0) {
        int acc = 0;
        for (int l = 0; l < 64; ++l) {
            off[l] = acc;
            acc += scal->level_count[l];
        }
        off[64] = acc;
        /* the levels must account for exactly the cells built (else the
         * level lists would not fit the table: stop, flag, write nothing) */
        consistent = acc == scal->ncell;
        if (!consistent) atomicOr(&scal->status, NBHIP_ERROR);
    }
    __syncthreads();
    if (blockIdx.x == 0 && tid <= 64) scal->level_off[tid] = off[tid];
    const int nn = (!consistent || (scal->status & (NBHIP_CELL_OVERFLOW_ERROR | NBHIP_ERROR)))
                       ? 0 : scal->nnode;
    for (int c0 = blockIdx.x * NBH_LS_CHUNK; c0 < nn; c0 += gridDim.x * NBH_LS_CHUNK) {
        const int cend = min(c0 + NBH_LS_CHUNK, nn);
        if (tid < 64) lh[tid] = 0;
        __syncthreads();
        for (int i = c0 + tid; i < cend; i += 256) {
            const int l = nlevel[i];
            if (l >= 0 && l <= 63) atomicAdd(&lh[l], 1);
        }
        __syncthreads();
        if (tid < 64) {
            base[tid] = lh[tid] ? atomicAdd(&scal->level_fill[tid], lh[tid]) : 0;
            lh[tid] = 0;
        }
        __syncthreads();
        for (int i = c0 + tid; i < cend; i += 256) {
            const int l = nlevel[i];
            if (l >= 0 && l <= 63) {
                const int r = atomicAdd(&lh[l], 1);
                level_list[off[l] + base[l] + r] = i;
            }
        }
        __syncthreads();
    }
}

/* ---------------------------------------------------------------------- */
/* bottom-up summarisation: block-local subtrees, then the spanning cells  */
/* ---------------------------------------------------------------------- */
/* hackCofM + findRCrit + nbCheckTreeStructure (nbody/src/nbody_tree.c:
 * 345-424) and hackQuad (:70-121), children in octant (pre-)order.
 *
 * Three launches, no host synchronisation and no cross-workgroup waits:
 *  - k_sum_blocks: workgroup b takes the massive bodies [b*NBH_SUM_BODIES,
 *    (b+1)*NBH_SUM_BODIES) and the pre-order range [P0, P1) of their leaves
 *    and of the cells whose first body they are.  A cell of that range whose
 *    subtree ends inside it (Next <= P1) is "inside": all its descendants
 *    are in the range, so the workgroup summarises its inside cells level by
 *    level, deepest first, with a barrier between levels.  Cells whose
 *    subtree runs past P1 ("spanning") all contain the block's last body, so
 *    there is at most one per level: they go to the block's own 64 slots.
 *  - k_sum_super: workgroup s takes NBH_SUM_GROUP blocks (a super-block) and
 *    their spanning cells; those whose subtree ends inside the super-block
 *    are summarised the same way (their children are inside cells, bodies
 *    or deeper cells of this list), the rest (again at most one per level)
 *    go to the super-block's 64 slots.
 *  - k_sum_top: one workgroup summarises the super-blocks' spanning cells
 *    (a few per level and super-block), deepest first.
 * A cell's children come from its child list (k_fill_kids): the pre-order
 * indices of its children, sorted into octant order and checked against the
 * Next links, then loaded together instead of one after another. */

/* The group walk's table (precision MIXED_GROUP): after the summaries,
 * every node's record goes to its parent's child block (slot nsim +
 * kbase[parent] + rank, rank in octant (= pre-order) order among the
 * siblings) and every root's to slot sim (k_group_table). */
struct GroupTab {
    NodeG* T;
    const unsigned* nkid;
    const unsigned* kbase;
    int nsim;
};

/* Group-walk record of node c (a finished cell or a body). */
__device__ __forceinline__ NodeG nbh_group_record(const NodeW& w, int c,
                                                  const double* __restrict__ nquad,
                                                  const GroupTab& g, double rc_scale, bool quad)
{
    NodeG r;
    r.x = w.x;
    r.y = w.y;
    r.z = w.z;
    r.m = (float) w.m;
    if (w.body >= 0) {
        r.rcrit2 = -1.0f;
        r.link = w.body;
        r.count = 0;
        for (int k = 0; k < 6; ++k) r.q[k] = 0.0f;
    } else {
        r.rcrit2 = __double2float_ru(w.rcrit2 * rc_scale);
        r.link = g.nsim + (int) g.kbase[c];
        r.count = __popc(g.nkid[c]);
        for (int k = 0; k < 6; ++k) r.q[k] = quad ? (float) nquad[6 * (size_t) c + k] : 0.0f;
    }
    return r;
}

/* Child lists: kid[kbase[p] .. kbase[p] + popcount(nkid[p])) holds the
 * children of cell p in octant order (nrank: a child's octant; nkid[p]: the
 * mask of occupied octants, both from the emission). */
__global__ __launch_bounds__(256) void k_fill_kids(const int* __restrict__ nparent,
                                                   const int* __restrict__ nrank,
                                                   const unsigned* __restrict__ nkid,
                                                   const unsigned* __restrict__ kbase,
                                                   int* __restrict__ kid, int node_cap,
                                                   DevScalars* __restrict__ scal)
{
    if (scal->status & (NBHIP_CELL_OVERFLOW_ERROR | NBHIP_ERROR)) return;
    const int nn = min(scal->nnode, node_cap);
    for (int q = blockIdx.x * 256 + threadIdx.x; q < nn; q += gridDim.x * 256) {
        const int par = nparent[q];
        if (par < 0) continue;
        const int r = nrank[q];
        const unsigned mask = nkid[par];
        if (par >= q || r < 0 || r > 7 || !((mask >> r) & 1u)) {
            atomicOr(&scal->status, NBHIP_ERROR);
            continue;
        }
        kid[kbase[par] + __popc(mask & ((1u << r) - 1u))] = q;
    }
}

struct SumView {
    NodeW* node;
    double* nquad;
    const double* nmid;
    const int* nnext;
    const int* nsim;
    const int* nlevel;
    const int* kid;
    SimTab* sims;
    NodeF* nodef;
    GroupTab g;
    nbhip_ctx ctx;
    double rc_scale;
    int nn;
};

#define NBH_NO_KID 0x7fffffff   /* an unused slot of a sorted child list */
#define NBH_KID_CELL 0x80000000u /* flag of a child list entry: the child is a cell */

/* The structure half of a cell's summary: its children, sorted into
 * pre-order (= octant order), into ch[8] (unused slots NBH_NO_KID; cells
 * flagged NBH_KID_CELL), checked against the Next links; false on a child
 * list that disagrees with them (a corrupt table).  Reads structure only
 * (levels, Next links, child lists), so it may run before the children are
 * summarised. */
__device__ __forceinline__ bool nbh_cell_children(const SumView& v, int p, int* __restrict__ out)
{
    /* without branches (a failed check clears ok and later loads use safe
     * indices), so that two inlined calls overlap their three rounds of
     * dependent loads */
    const int level = v.nlevel[p];
    const int end = v.nnext[p];
    const int nk = __popc(v.g.nkid[p]);
    const int kbr = (int) v.g.kbase[p];
    bool ok = !(level < 0 || end <= p + 1 || end > v.nn || nk < 1 || nk > 8);
    const int kb = ok ? kbr : 0;
    int ch[8];
    #pragma unroll
    for (int j = 0; j < 8; ++j) ch[j] = (ok && j < nk) ? v.kid[kb + j] : NBH_NO_KID;
    /* Batcher's odd-even merge network for 8 keys (19 compare-exchanges,
     * static indices: the list stays in registers) */
#define NBH_CE(a, b)                                                    \
    do {                                                                \
        const int lo_ = min(ch[a], ch[b]), hi_ = max(ch[a], ch[b]);     \
        ch[a] = lo_;                                                    \
        ch[b] = hi_;                                                    \
    } while (0)
    NBH_CE(0, 1); NBH_CE(2, 3); NBH_CE(4, 5); NBH_CE(6, 7);
    NBH_CE(0, 2); NBH_CE(1, 3); NBH_CE(4, 6); NBH_CE(5, 7);
    NBH_CE(1, 2); NBH_CE(5, 6);
    NBH_CE(0, 4); NBH_CE(1, 5); NBH_CE(2, 6); NBH_CE(3, 7);
    NBH_CE(2, 4); NBH_CE(3, 5);
    NBH_CE(1, 2); NBH_CE(3, 4); NBH_CE(5, 6);
#undef NBH_CE
    ok = ok && ch[0] == p + 1;
    #pragma unroll
    for (int j = 0; j < 8; ++j)
        if (j < nk) ok = ok && ch[j] >= 0 && ch[j] < v.nn;
    int lv[8], nx[8];
    #pragma unroll
    for (int j = 0; j < 8; ++j) {
        const bool use = ok && j < nk;
        const int c = use ? ch[j] : 0;
        lv[j] = use ? v.nlevel[c] : -1;
        const int nn = use ? v.nnext[c] : 0;
        nx[j] = use ? (lv[j] < 0 ? ch[j] + 1 : nn) : 0;
    }
    /* the list must be the Next chain p+1 -> ... -> end */
    #pragma unroll
    for (int j = 0; j < 8; ++j)
        if (j < nk) ok = ok && nx[j] == (j + 1 < nk ? ch[j + 1] : end);
    #pragma unroll
    for (int j = 0; j < 8; ++j)
        out[j] = j < nk ? (int) ((unsigned) ch[j] | (lv[j] >= 0 ? NBH_KID_CELL : 0u)) : NBH_NO_KID;
    return ok;
}

/* The value half: the cell's record (centre of mass, mass, critical
 * radius, quadrupole, and the mixed walk's record) from its children's
 * finished records; chl is the cell's checked, sorted, flagged child list.
 * The children's positions and masses and the quadrupoles of the child
 * cells are loaded together (the flags say which children are cells). */
__device__ __forceinline__ void nbh_cell_summary(const SumView& v, int p, const int* __restrict__ chl)
{
    NodeW* __restrict__ node = v.node;
    int ch[8];
    bool cell[8];
    int nk = 0;
    #pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int e = chl[j];
        nk += e != NBH_NO_KID;
        cell[j] = e != NBH_NO_KID && ((unsigned) e & NBH_KID_CELL);
        ch[j] = e & 0x7fffffff;
    }
    const int level = v.nlevel[p];
    const int sim = v.nsim[p];
    const double rsize = v.sims[sim].rsize;
    const int end = v.nnext[p];
    const bool quad = v.ctx.use_quad != 0;
    double4 w[8];                                  /* x y z m */
    double cq[8][6];
    #pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (j < nk) w[j] = *reinterpret_cast<const double4*>(&node[ch[j]]);
        #pragma unroll
        for (int k = 0; k < 6; ++k)
            cq[j][k] = (quad && j < nk && cell[j]) ? v.nquad[6 * (size_t) ch[j] + k] : 0.0;
    }

    double mass = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    #pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (j < nk) {
            mass += w[j].w;
            cx += w[j].w * w[j].x;
            cy += w[j].w * w[j].y;
            cz += w[j].w * w[j].z;
        }
    }
    const double mx = v.nmid[3 * p + 0], my = v.nmid[3 * p + 1], mz = v.nmid[3 * p + 2];
    if (mass > 0.0) {
        cx /= mass;
        cy /= mass;
        cz /= mass;
    } else {
        cx = mx; cy = my; cz = mz;
    }
    const double psize = ldexp(rsize, -level);
    const double half = 0.5 * psize;
    if (cx < mx - half || cx > mx + half || cy < my - half || cy > my + half ||
        cz < mz - half || cz > mz + half)
        atomicOr(&v.sims[sim].flags, NBHIP_TREE_STRUCTURE_ERROR);
    const double rc2 = nbh_rcrit2(v.ctx, mx, my, mz, rsize, cx, cy, cz, psize);

    double* pq = &v.nquad[6 * (size_t) p];
    if (quad) {
        double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0, q4 = 0.0, q5 = 0.0;
        #pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (j >= nk) continue;
            const double dx = w[j].x - cx, dy = w[j].y - cy, dz = w[j].z - cz;
            const double drsq = dz * dz + (dy * dy + dx * dx);
            const double m = w[j].w;
            double e0 = m * (3.0 * (dx * dx) - drsq);
            double e1 = m * (3.0 * (dx * dy));
            double e2 = m * (3.0 * (dx * dz));
            double e3 = m * (3.0 * (dy * dy) - drsq);
            double e4 = m * (3.0 * (dy * dz));
            double e5 = m * (3.0 * (dz * dz) - drsq);
            if (cell[j]) {
                e0 += cq[j][0]; e1 += cq[j][1]; e2 += cq[j][2];
                e3 += cq[j][3]; e4 += cq[j][4]; e5 += cq[j][5];
            }
            q0 += e0; q1 += e1; q2 += e2; q3 += e3; q4 += e4; q5 += e5;
        }
        pq[0] = q0; pq[1] = q1; pq[2] = q2; pq[3] = q3; pq[4] = q4; pq[5] = q5;
        if (v.nodef) {
            NodeF& f = v.nodef[p];
            f.q[0] = (float) q0; f.q[1] = (float) q1; f.q[2] = (float) q2;
            f.q[3] = (float) q3; f.q[4] = (float) q4; f.q[5] = (float) q5;
        }
    } else {
        pq[0] = pq[1] = pq[2] = pq[3] = pq[4] = pq[5] = 0.0;
        if (v.nodef)
            for (int k = 0; k < 6; ++k) v.nodef[p].q[k] = 0.0f;
    }
    NodeW r;
    r.x = cx;
    r.y = cy;
    r.z = cz;
    r.m = mass;
    r.rcrit2 = rc2;
    r.next = end;
    r.body = -1;
    node[p] = r;
    if (v.nodef) {
        /* the mixed walk's record: FP64 position, the rest in FP32 */
        NodeF& f = v.nodef[p];
        f.x = cx;
        f.y = cy;
        f.z = cz;
        f.m = (float) mass;
        /* the coherent walk may open more (rc_scale >= 1; 1 otherwise) */
        f.rcrit2 = (float) (rc2 * v.rc_scale);
        f.next = end;
        f.body = -1;
    }
}

/* One cell's summary; false on a child list that disagrees with the Next
 * links (a corrupt table). */
__device__ __forceinline__ bool nbh_summarize_cell(const SumView& v, int p)
{
    int ch[8];
    if (!nbh_cell_children(v, p, ch)) return false;
    nbh_cell_summary(v, p, ch);
    return true;
}

#define NBH_SUM_BODIES 1024    /* massive bodies per k_sum_blocks workgroup */
#define NBH_SUM_THREADS 256
#define NBH_SUMB_THREADS 128   /* k_sum_blocks: two waves, so that with its ~220 VGPRs
                                  four workgroups share a CU and all blocks of a
                                  10^6-body tree run at once */
#define NBH_SUM_GROUP 64       /* blocks per k_sum_super workgroup */
#define NBH_SPAN_SLOTS 64      /* spanning cells one block can hand up (one per level) */
#define NBH_SPAN_THREADS 512
#define NBH_SUM_CAP 1024       /* cells whose child lists a workgroup keeps in LDS */
#define NBH_SUMB_REG 16        /* k_sum_blocks: table entries per thread held in registers */

/* off[l] = cells of levels < l, off[64] = all (exclusive scan of the 64
 * level counts cnt[], by the first wave; callers synchronise after) */
__device__ __forceinline__ void nbh_level_offsets(const int* cnt, int* off)
{
    if (threadIdx.x < 64) {
        const int c = cnt[threadIdx.x];
        int x = c;
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(x, o, 64);
            if ((int) threadIdx.x >= o) x += y;
        }
        off[threadIdx.x] = x - c;
        if (threadIdx.x == 63) off[64] = x;
    }
}

/* Summaries of the cells list[0 .. off[64]), grouped by level (level l at
 * list[off[l] .. off[l+1])), deepest level first, one barrier per level.
 * Up to NBH_SUM_CAP cells, every cell's child list is read, sorted and
 * checked first, all levels at once, and kept in LDS (s_ch: two int4 per
 * cell), so that a level then waits on one round of loads -- its children's
 * records -- instead of the chain nlevel/kbase -> kid -> children.  More
 * cells: each level does both halves.  False on a corrupt child list. */
template <int NT>
__device__ bool nbh_sum_levels(const SumView& v, const int* list, const int* off, int4* s_ch)
{
    const int tid = threadIdx.x;
    const int total = off[64];
    bool ok = true;
    if (total <= NBH_SUM_CAP) {
        /* two cells per round, their loads overlapped */
        auto keep = [&](int i, bool good, const int* ch) {
            if (good) {
                s_ch[2 * i] = make_int4(ch[0], ch[1], ch[2], ch[3]);
                s_ch[2 * i + 1] = make_int4(ch[4], ch[5], ch[6], ch[7]);
            } else {
                ok = false;                               /* marked: a list never starts empty */
                s_ch[2 * i] = make_int4(NBH_NO_KID, NBH_NO_KID, NBH_NO_KID, NBH_NO_KID);
            }
        };
        for (int i = tid; i < total; i += 2 * NT) {
            const int i1 = i + NT;
            const bool has1 = i1 < total;
            int c0[8], c1[8];
            const bool g0 = nbh_cell_children(v, list[i], c0);
            const bool g1 = nbh_cell_children(v, list[has1 ? i1 : i], c1);
            keep(i, g0, c0);
            if (has1) keep(i1, g1, c1);
        }
        __syncthreads();
        for (int l = 63; l >= 0; --l) {
            if (off[l + 1] == off[l]) continue;           /* uniform */
            for (int i = off[l] + tid; i < off[l + 1]; i += NT) {
                const int4 a = s_ch[2 * i];
                if (a.x == NBH_NO_KID) continue;
                const int4 b = s_ch[2 * i + 1];
                const int ch[8] = { a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w };
                nbh_cell_summary(v, list[i], ch);
            }
            __syncthreads();
        }
    } else {
        for (int l = 63; l >= 0; --l) {
            if (off[l + 1] == off[l]) continue;           /* uniform */
            for (int i = off[l] + tid; i < off[l + 1]; i += NT)
                ok = nbh_summarize_cell(v, list[i]) && ok;
            __syncthreads();
        }
    }
    return ok;
}

/* Levels holding cells in [lo, hi): counts into cnt[64] (LDS), then an
 * exclusive scan into off[65]; callers add their own cells. */
__global__ __launch_bounds__(NBH_SUMB_THREADS) void k_sum_blocks(SumView v,
                                                                const unsigned* __restrict__ cincl,
                                                                int* __restrict__ blist,
                                                                int* __restrict__ span,
                                                                int* __restrict__ span_n,
                                                                DevScalars* __restrict__ scal)
{
    __shared__ int cnt[64];
    __shared__ int off[65];
    __shared__ int fill[64];
    __shared__ int nspan;
    __shared__ int s_list[NBH_SUM_CAP];
    __shared__ int4 s_ch[2 * NBH_SUM_CAP];
    const int tid = threadIdx.x;
    if (scal->status & (NBHIP_CELL_OVERFLOW_ERROR | NBHIP_ERROR)) return;
    const int M = scal->nmassive;
    const int b0 = blockIdx.x * NBH_SUM_BODIES;
    if (b0 >= M) return;
    const int b1 = min(b0 + NBH_SUM_BODIES, M);
    const int P0 = b0 + (b0 > 0 ? (int) cincl[b0 - 1] : 0);
    const int P1 = min((b1 - 1) + (int) cincl[b1 - 1] + 1, v.nn);
    if (tid < 64) {
        cnt[tid] = 0;
        fill[tid] = 0;
    }
    if (tid == 0) nspan = 0;
    __syncthreads();
    /* the range's levels and Next links, loaded once and all at once (up to
     * NBH_SUMB_REG entries per thread; a longer range loads them per pass) */
    constexpr int REG = NBH_SUMB_REG;
    const bool in_regs = P1 - P0 <= REG * NBH_SUMB_THREADS;
    int rl[REG], rn[REG];
    #pragma unroll
    for (int k = 0; k < REG; ++k) {
        const int q = P0 + tid + k * NBH_SUMB_THREADS;
        const bool h = in_regs && q < P1;
        rl[k] = h ? v.nlevel[q] : -1;
        rn[k] = h ? v.nnext[q] : 0;
    }
    /* inside cells per level; spanning cells to the second launch */
    auto classify = [&](int q, int l, int nx) {
        if (l < 0) return;                        /* a body's leaf */
        if (l > 63) {
            atomicOr(&scal->status, NBHIP_ERROR);
            return;
        }
        if (nx <= P1) {
            atomicAdd(&cnt[l], 1);
        } else {
            const int k = atomicAdd(&nspan, 1);
            if (k < NBH_SPAN_SLOTS)
                span[(size_t) blockIdx.x * NBH_SPAN_SLOTS + k] = q;
            else
                atomicOr(&scal->status, NBHIP_ERROR);   /* not nested: a corrupt table */
        }
    };
    if (in_regs) {
        #pragma unroll
        for (int k = 0; k < REG; ++k) classify(P0 + tid + k * NBH_SUMB_THREADS, rl[k], rn[k]);
    } else {
        for (int q = P0 + tid; q < P1; q += NBH_SUMB_THREADS) classify(q, v.nlevel[q], v.nnext[q]);
    }
    __syncthreads();
    if (tid == 0) span_n[blockIdx.x] = min(nspan, NBH_SPAN_SLOTS);
    nbh_level_offsets(cnt, off);
    __syncthreads();
    /* the block's inside cells grouped by level: in LDS, or, past
     * NBH_SUM_CAP, in the block's own slice [P0, P0 + count) of blist
     * (slices of different blocks are disjoint) */
    int* lst = off[64] <= NBH_SUM_CAP ? s_list : blist + P0;
    auto place = [&](int q, int l, int nx) {
        if (l < 0 || l > 63 || nx > P1) return;
        lst[off[l] + atomicAdd(&fill[l], 1)] = q;
    };
    if (in_regs) {
        #pragma unroll
        for (int k = 0; k < REG; ++k) place(P0 + tid + k * NBH_SUMB_THREADS, rl[k], rn[k]);
    } else {
        for (int q = P0 + tid; q < P1; q += NBH_SUMB_THREADS) place(q, v.nlevel[q], v.nnext[q]);
    }
    __syncthreads();
    if (!nbh_sum_levels<NBH_SUMB_THREADS>(v, lst, off, s_ch)) atomicOr(&scal->status, NBHIP_ERROR);
}

/* Cells handed up by several lower workgroups (cand[0..nc) in LDS or
 * global memory), grouped by level in LDS and summarised deepest level
 * first; those whose subtree ends past P1 go to out[0..NBH_SPAN_SLOTS)
 * instead (out may be null when P1 covers the whole table). */
template <int NT>
__device__ void nbh_sum_handed(const SumView& v, const int* cand, int nc, int P1, int* out,
                               int* out_n, int* list, int* cnt, int* off, int* fill, int* nspan,
                               int4* s_ch, DevScalars* scal)
{
    const int tid = threadIdx.x;
    if (tid < 64) {
        cnt[tid] = 0;
        fill[tid] = 0;
    }
    if (tid == 0) *nspan = 0;
    __syncthreads();
    for (int i = tid; i < nc; i += NT) {
        const int q = cand[i];
        const int l = v.nlevel[q];
        if (l < 0 || l > 63) {
            atomicOr(&scal->status, NBHIP_ERROR);
            continue;
        }
        if (v.nnext[q] <= P1) {
            atomicAdd(&cnt[l], 1);
        } else {
            const int k = atomicAdd(nspan, 1);
            if (out && k < NBH_SPAN_SLOTS) out[k] = q;
            else atomicOr(&scal->status, NBHIP_ERROR);
        }
    }
    __syncthreads();
    nbh_level_offsets(cnt, off);
    if (tid == 0 && out_n) *out_n = min(*nspan, NBH_SPAN_SLOTS);
    __syncthreads();
    for (int i = tid; i < nc; i += NT) {
        const int q = cand[i];
        const int l = v.nlevel[q];
        if (l < 0 || l > 63 || v.nnext[q] > P1) continue;
        list[off[l] + atomicAdd(&fill[l], 1)] = q;
    }
    __syncthreads();
    if (!nbh_sum_levels<NT>(v, list, off, s_ch)) atomicOr(&scal->status, NBHIP_ERROR);
}

/* Workgroup s: the spanning cells of blocks [s*NBH_SUM_GROUP, ...). */
__global__ __launch_bounds__(NBH_SUM_THREADS) void k_sum_super(SumView v,
                                                               const unsigned* __restrict__ cincl,
                                                               const int* __restrict__ span,
                                                               const int* __restrict__ span_n,
                                                               int* __restrict__ span2,
                                                               int* __restrict__ span2_n,
                                                               DevScalars* __restrict__ scal)
{
    __shared__ int cand[NBH_SUM_GROUP * NBH_SPAN_SLOTS];
    __shared__ int list[NBH_SUM_GROUP * NBH_SPAN_SLOTS];
    __shared__ int cnt[64];
    __shared__ int off[65];
    __shared__ int fill[64];
    __shared__ int nspan;
    __shared__ int ncand;
    __shared__ int4 s_ch[2 * NBH_SUM_CAP];
    const int tid = threadIdx.x;
    if (scal->status & (NBHIP_CELL_OVERFLOW_ERROR | NBHIP_ERROR)) return;
    const int M = scal->nmassive;
    const int nblk = (M + NBH_SUM_BODIES - 1) / NBH_SUM_BODIES;
    const int k0 = blockIdx.x * NBH_SUM_GROUP;
    if (k0 >= nblk) return;
    const int k1 = min(k0 + NBH_SUM_GROUP, nblk);
    const int b1 = min(k1 * NBH_SUM_BODIES, M);
    const int P1 = min((b1 - 1) + (int) cincl[b1 - 1] + 1, v.nn);
    if (tid == 0) ncand = 0;
    __syncthreads();
    for (int k = k0 + tid; k < k1; k += NBH_SUM_THREADS) {
        const int c = span_n[k];
        const int base = atomicAdd(&ncand, c);
        for (int j = 0; j < c; ++j) cand[base + j] = span[(size_t) k * NBH_SPAN_SLOTS + j];
    }
    __syncthreads();
    nbh_sum_handed<NBH_SUM_THREADS>(v, cand, ncand, P1, span2 + (size_t) blockIdx.x * NBH_SPAN_SLOTS,
                                    span2_n + blockIdx.x, list, cnt, off, fill, &nspan, s_ch,
                                    scal);
}

/* One workgroup: every super-block's spanning cells. */
__global__ __launch_bounds__(NBH_SPAN_THREADS) void k_sum_top(SumView v,
                                                              const int* __restrict__ span2,
                                                              const int* __restrict__ span2_n,
                                                              int nsuper, int* __restrict__ cand,
                                                              int* __restrict__ list,
                                                              DevScalars* __restrict__ scal)
{
    __shared__ int cnt[64];
    __shared__ int off[65];
    __shared__ int fill[64];
    __shared__ int nspan;
    __shared__ int ncand;
    __shared__ int4 s_ch[2 * NBH_SUM_CAP];
    const int tid = threadIdx.x;
    if (scal->status & (NBHIP_CELL_OVERFLOW_ERROR | NBHIP_ERROR)) return;
    const int M = scal->nmassive;
    const int nblk = (M + NBH_SUM_BODIES - 1) / NBH_SUM_BODIES;
    const int ns = min(nsuper, (nblk + NBH_SUM_GROUP - 1) / NBH_SUM_GROUP);
    if (tid == 0) ncand = 0;
    __syncthreads();
    for (int k = tid; k < ns; k += NBH_SPAN_THREADS) {
        const int c = span2_n[k];
        const int base = atomicAdd(&ncand, c);
        for (int j = 0; j < c; ++j) cand[base + j] = span2[(size_t) k * NBH_SPAN_SLOTS + j];
    }
    __syncthreads();
    nbh_sum_handed<NBH_SPAN_THREADS>(v, cand, ncand, v.nn, nullptr, nullptr, list, cnt, off,
                                     fill, &nspan, s_ch, scal);
}

/* The group walk's table from the finished summaries: one thread per node
 * (no level order needed: every record is final). */
__global__ __launch_bounds__(256) void k_group_table(const NodeW* __restrict__ node,
                                                     const double* __restrict__ nquad,
                                                     const int* __restrict__ nparent,
                                                     const int* __restrict__ nsim,
                                                     const int* __restrict__ nrank, GroupTab g,
                                                     double rc_scale, int use_quad, int node_cap,
                                                     DevScalars* __restrict__ scal)
{
    if (scal->status & (NBHIP_CELL_OVERFLOW_ERROR | NBHIP_ERROR)) return;
    const int nn = min(scal->nnode, node_cap);
    for (int q = blockIdx.x * 256 + threadIdx.x; q < nn; q += gridDim.x * 256) {
        const NodeW w = node[q];
        const int par = nparent[q];
        const NodeG r = nbh_group_record(w, q, nquad, g, rc_scale, use_quad != 0);
        if (par < 0) {
            g.T[nsim[q]] = r;
            continue;
        }
        /* the rank among the siblings: the occupied octants below its own
         * (the child list's order; k_fill_kids has checked the slot) */
        const int o = nrank[q] & 7;
        g.T[g.nsim + g.kbase[par] + __popc(g.nkid[par] & ((1u << o) - 1u))] = r;
    }
}

/* End of a KDK step (nbRunSystemPlain's loop, nbody/src/nbody_plain.c:
 * 409-448): counts the step and, when a simulation's tree or walk reported
 * a fatal status (a structure, depth or internal error, or incest without
 * allowIncest), halts the state so that the remaining steps of a replayed
 * graph do nothing -- the run stops at the step that failed, as the
 * reference's loop returns there. */
__global__ __launch_bounds__(256) void k_step_end(int nsim, const SimTab* __restrict__ sims,
                                                  int allow_incest, DevScalars* __restrict__ scal)
{
    __shared__ int fatal;
    if (scal->halt) return;
    if (threadIdx.x == 0) fatal = (scal->status & ~NBHIP_STATUS_ALL_WARNINGS) != 0u;
    __syncthreads();
    for (int k = threadIdx.x; k < nsim; k += 256) {
        const SimTab& t = sims[k];
        if ((t.flags & ~NBHIP_STATUS_ALL_WARNINGS) || (t.incest && !allow_incest)) fatal = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        scal->steps_done += 1u;
        if (fatal) scal->halt = 1;
    }
}

/* ---------------------------------------------------------------------- */
/* host launchers                                                          */
/* ---------------------------------------------------------------------- */
static inline int nbh_blocks(int n, int per) { return n > 0 ? (n + per - 1) / per : 1; }

void nbh_bbox(nbhip_state* st, int drift)
{
    const double dt = st->ctx.timestep;
    hipLaunchKernelGGL(k_drift_bbox, dim3(nbh_blocks(st->n, NBH_BBOX_CHUNK)), dim3(256), 0,
                       st->stream,
                       st->n, drift, dt, 0.5 * dt, st->px, st->py, st->pz, st->vx, st->vy, st->vz,
                       st->ax, st->ay, st->az, st->m, st->pos4, st->simid, st->sims, st->scal);
    hipLaunchKernelGGL(k_rsize, dim3(1), dim3(256), 0, st->stream, st->nsim, st->sims, st->scal);
}

void nbh_step_end(nbhip_state* st)
{
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(256), 0, st->stream, st->nsim, st->sims,
                       st->ctx.allow_incest, st->scal);
}

static void nbh_mark(nbhip_state* st, int k)
{
    if (st->profiling) (void) hipEventRecord(st->ev[k], st->stream);
}

/* graph_safe: no host synchronisation (all 64 level launches, grid-stride) */
/* Level lists of the current tree (cells grouped by level, in level_list
 * with scal->level_off) and their counts on the host: the adjoint's
 * top-down pass over the cells needs them; the summaries do not. */
__global__ void k_level_fill_reset(DevScalars* __restrict__ scal)
{
    if (threadIdx.x < 64) {
        scal->level_fill[threadIdx.x] = 0;
        scal->level_count[threadIdx.x] = 0;
    }
}

/* Cells per level of the current tree (grid-stride over the nodes built). */
__global__ __launch_bounds__(256) void k_level_count(const int* __restrict__ nlevel,
                                                     DevScalars* __restrict__ scal)
{
    __shared__ int lh[64];
    const int tid = threadIdx.x;
    if (tid < 64) lh[tid] = 0;
    __syncthreads();
    const int nn = (scal->status & NBHIP_CELL_OVERFLOW_ERROR) ? 0 : scal->nnode;
    for (int i = blockIdx.x * 256 + tid; i < nn; i += gridDim.x * 256) {
        const int l = nlevel[i];
        if (l >= 0 && l <= 63) atomicAdd(&lh[l], 1);
    }
    __syncthreads();
    if (tid < 64 && lh[tid]) atomicAdd(&scal->level_count[tid], lh[tid]);
}

unsigned nbh_level_lists(nbhip_state* st)
{
    hipStream_t s = st->stream;
    hipLaunchKernelGGL(k_level_fill_reset, dim3(1), dim3(64), 0, s, st->scal);
    hipLaunchKernelGGL(k_level_count, dim3(nbh_blocks(2 * st->n + 1024, NBH_LS_CHUNK)), dim3(256),
                       0, s, st->nlevel, st->scal);
    hipLaunchKernelGGL(k_level_scatter, dim3(nbh_blocks(2 * st->n + 1024, NBH_LS_CHUNK)),
                       dim3(256), 0, s, st->nlevel, st->level_list, st->scal);
    hipError_t e = hipMemcpyAsync(st->h_counts_pinned, st->scal->level_count, 64 * sizeof(int),
                                  hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        nbh_set_error("level lists: %s", hipGetErrorString(e));
        return NBHIP_GPU_ERROR;
    }
    memcpy(st->h_level_count, st->h_counts_pinned, sizeof(st->h_level_count));
    st->h_levels_valid = 1;
    return NBHIP_SUCCESS;
}

void nbh_build(nbhip_state* st, int graph_safe)
{
    (void) graph_safe;             /* no step of the build reads anything back */
    const int n = st->n;
    const int g256 = nbh_blocks(n, 256);
    hipStream_t s = st->stream;

    hipLaunchKernelGGL(k_keys, dim3(g256), dim3(256), 0, s, n, st->px, st->py, st->pz, st->m,
                       st->simid, st->sims, st->sim_shift, st->top_levels, st->key, st->perm);
    unsigned long long* kin = st->key;
    unsigned long long* kout = st->key_alt;
    unsigned* vin = st->perm;
    unsigned* vout = st->perm_alt;
    const int nb = st->sort_blocks;
    for (int pass = 0; pass < 8; ++pass) {
        const int shift = 8 * pass;
        hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(256), 0, s, n, shift, kin, st->hist, nb);
        hipLaunchKernelGGL(k_radix_scan, dim3(256), dim3(256), 0, s, st->hist, st->dtot, nb);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(256), 0, s, n, shift, kin, vin, kout,
                           vout, st->hist, st->dtot, nb);
        unsigned long long* tk = kin; kin = kout; kout = tk;
        unsigned* tv = vin; vin = vout; vout = tv;
    }
    /* 8 passes: sorted data is back in st->key / st->perm */
    hipLaunchKernelGGL(k_count_massive, dim3(g256), dim3(256), 0, s, n, st->key, st->scal);
    /* perm_alt is free after the eighth pass */
    hipLaunchKernelGGL(k_tiefix_rank, dim3(g256), dim3(256), 0, s, st->key, st->perm, st->perm_alt,
                       st->px, st->py, st->pz, st->simid, st->sims, st->top_levels, st->scal);
    hipLaunchKernelGGL(k_tiefix_copy, dim3(g256), dim3(256), 0, s, st->key, st->perm,
                       st->perm_alt, st->scal);
    nbh_mark(st, 2);

    hipLaunchKernelGGL(k_gather, dim3(g256), dim3(256), 0, s, n, st->perm, st->pos4, st->key,
                       st->sim_shift, st->simid, st->sims, st->spos, st->simid_s, st->su,
                       st->scal);
    hipLaunchKernelGGL(k_delta_count, dim3(g256), dim3(256), 0, s, n, st->su, st->simid_s,
                       st->dadj, st->cincl, st->sims, st->scal);
    const int nscan = nbh_blocks(n, NBH_SCAN_TILE);
    hipLaunchKernelGGL(k_scan_tiles, dim3(nscan), dim3(1024), 0, s, st->cincl, st->scan_part,
                       st->scal);
    hipLaunchKernelGGL(k_scan_part, dim3(1), dim3(1024), 0, s, st->scan_part, st->scal);
    hipLaunchKernelGGL(k_scan_add, dim3(nscan), dim3(1024), 0, s, st->cincl, st->scan_part,
                       st->scal);
    hipLaunchKernelGGL(k_sim_roots, dim3(g256), dim3(256), 0, s, st->simid_s, st->cincl,
                       st->node_cap, st->sims, st->nkid, st->scal);
    hipLaunchKernelGGL(k_sim_ends, dim3(1), dim3(256), 0, s, st->nsim, st->sims, st->scal);
    hipLaunchKernelGGL(k_dhier_top, dim3(1), dim3(1024), 0, s, n, st->dadj, st->scal);
    hipLaunchKernelGGL(k_emit_cells, dim3(g256), dim3(256), 0, s, n, st->dadj, st->cincl,
                       st->su, st->simid_s, st->nmid, st->nlevel, st->nparent, st->nsim_node,
                       st->sims, st->nkid, st->nrank, st->nnext, st->scal);
    hipLaunchKernelGGL(k_emit_leaves, dim3(g256), dim3(256), 0, s, n, st->dadj, st->su, st->cincl,
                       st->spos, st->simid_s, st->node, st->nmid, st->nlevel, st->nparent,
                       st->nsim_node, st->sims, st->nodef, st->nkid, st->nrank, st->nnext,
                       st->scal);
    /* each cell's child block (kbase) and child list (kid) */
    const int nt = nbh_blocks(st->node_cap, NBH_SCAN_TILE);
    hipLaunchKernelGGL(k_nscan_tiles, dim3(nt), dim3(1024), 0, s, st->nkid, st->kbase,
                       st->scan_part, st->node_cap, st->scal);
    hipLaunchKernelGGL(k_nscan_part, dim3(1), dim3(1024), 0, s, st->scan_part, st->node_cap,
                       st->scal);
    hipLaunchKernelGGL(k_nscan_add, dim3(nt), dim3(1024), 0, s, st->kbase, st->scan_part,
                       st->node_cap, st->scal);
    hipLaunchKernelGGL(k_fill_kids, dim3(nbh_blocks(st->node_cap, 256 * 8)), dim3(256), 0, s,
                       st->nparent, st->nrank, st->nkid, st->kbase, st->kid, st->node_cap,
                       st->scal);
    nbh_mark(st, 3);
    /* level lists are built only for the adjoint's top-down pass (on its
     * request: nbh_level_lists) */
    st->h_levels_valid = 0;
    SumView v;
    v.node = st->node;
    v.nquad = st->nquad;
    v.nmid = st->nmid;
    v.nnext = st->nnext;
    v.nsim = st->nsim_node;
    v.nlevel = st->nlevel;
    v.kid = st->kid;
    v.sims = st->sims;
    v.nodef = st->nodef;
    v.g = GroupTab{ nullptr, st->nkid, st->kbase, st->nsim };   /* table: k_group_table */
    v.ctx = st->ctx;
    v.rc_scale = st->open_scale;
    v.nn = st->node_cap;
    /* span: each block's spanning cells (NBH_SPAN_SLOTS each); span_sorted:
     * each super-block's, then the per-block and per-super-block counts,
     * then the top launch's candidates and level-sorted list */
    const int nblk = nbh_blocks(n, NBH_SUM_BODIES);
    const int nsup = nbh_blocks(nblk, NBH_SUM_GROUP);
    int* span2 = st->span_sorted;
    int* span_n = span2 + (size_t) nsup * NBH_SPAN_SLOTS;
    int* span2_n = span_n + nblk;
    int* top_cand = span2_n + nsup;
    int* top_list = top_cand + (size_t) nsup * NBH_SPAN_SLOTS;
    hipLaunchKernelGGL(k_sum_blocks, dim3(nblk), dim3(NBH_SUMB_THREADS), 0, s, v, st->cincl,
                       st->blist, st->span, span_n, st->scal);
    hipLaunchKernelGGL(k_sum_super, dim3(nsup), dim3(NBH_SUM_THREADS), 0, s, v, st->cincl,
                       st->span, span_n, span2, span2_n, st->scal);
    hipLaunchKernelGGL(k_sum_top, dim3(1), dim3(NBH_SPAN_THREADS), 0, s, v, span2, span2_n, nsup,
                       top_cand, top_list, st->scal);
    if (st->nodeg) {
        const GroupTab g{ st->nodeg, st->nkid, st->kbase, st->nsim };
        hipLaunchKernelGGL(k_group_table, dim3(nbh_blocks(st->node_cap, 256 * 4)), dim3(256), 0,
                           s, st->node, st->nquad, st->nparent, st->nsim_node, st->nrank, g,
                           st->open_scale, st->ctx.use_quad, st->node_cap, st->scal);
    }
    nbh_mark(st, 4);
}
