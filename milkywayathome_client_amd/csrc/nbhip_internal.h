/*
 * Internal declarations of the nbhip library (not part of the C ABI).
 *
 * HBM layout of one simulation (N bodies, node capacity C):
 *   bodies, original order, SoA FP64 : px py pz vx vy vz m ax ay az   (80 B/body)
 *   sort: 64-bit Morton word (levels 1-21) + 32-bit body index, ping-pong
 *   sorted bodies: spos (x y z m, 32 B) and 63-bit fixed-point coordinates
 *     su (3 x 8 B) of the massive bodies, used for the 189-bit Morton prefix
 *   adjacent prefixes dadj of the sorted massive bodies and their
 *     min-hierarchy (the binary radix tree, Karras 2012, is never stored:
 *     it is the Cartesian tree of dadj, found by nearest-smaller searches)
 *   node table in pre-order (the reference's More/Next threading order):
 *     NodeW (48 B: centre of mass, mass, rcrit2, Next index, body) + quad
 *     (48 B) + build-only level/parent/child counters/midpoint.
 */
#ifndef NBHIP_INTERNAL_H
#define NBHIP_INTERNAL_H

#include <hip/hip_runtime.h>
#include "../../include/nbhip.h"

/* Walk record of one node.  body >= 0: a massive body (its sorted index);
 * body < 0: a cell.  next: pre-order index of Next(node). */
struct __attribute__((aligned(16))) NodeW {
    double x, y, z, m;
    double rcrit2;
    int next;
    int body;
};

/* Walk record of the mixed-precision walk (64 B, one scalar load).
 * The position stays FP64: the walk forms each displacement as one FP64
 * subtraction rounded once to FP32, exact to FP32 however far the record and
 * the body are from the origin.  rcrit2 is -1 for bodies (always accepted);
 * q[6] is the quadrupole in FP32 (zero for bodies). */
struct __attribute__((aligned(16))) NodeF {
    double x, y, z;
    float m, rcrit2;
    int next, body;
    float q[6];
};

/* Record of the group walk's table, in which every cell's children occupy
 * one contiguous block (slots 0 .. nsim-1 hold the simulations' roots).
 * Position FP64 (displacements are formed relative to the group in FP64 and
 * rounded once); rcrit2 is the squared critical radius times the open scale,
 * rounded towards +infinity to FP32 (so the group test never accepts what
 * the FP64 criterion opens), -1 for bodies.  Cells: link = first child slot,
 * count = children (1..8); bodies: link = sorted body index, count = 0. */
struct __attribute__((aligned(16))) NodeG {
    double x, y, z;
    float m, rcrit2;
    int link, count;
    float q[6];
};

/* Exact criterion: the sources of a simulation of n bodies are summed in
 * this many consecutive segments (one per grid row), 1 below 4096 bodies. */
__host__ __device__ inline int nbh_exact_segments(int n)
{
    const int s = n / 2048;
    return s < 1 ? 1 : (s > 16 ? 16 : s);
}

/* Per-simulation device record (one per dwarf of a batch; one for a single
 * simulation). */
struct SimTab {
    double rsize;                  /* st->tree.rsize, persists, grows only */
    double key_scale;              /* 2^63 / rsize */
    double eps2;
    unsigned long long xyzmax_bits;/* max |x_i| of the last bbox pass */
    int body_off, body_n;          /* bodies, original order */
    int first;                     /* first sorted massive body, -1 if none */
    int root, end;                 /* pre-order segment [root, end), -1 if none */
    int maxdepth;
    unsigned flags;                /* sticky NBHIP_* bits of this simulation */
    unsigned incest;               /* a body of this simulation missed itself */
};

/* Device-resident scalars shared by all simulations of a state. */
struct DevScalars {
    int nmassive;                  /* massive bodies, all simulations */
    int ncell;
    int nnode;
    int maxdepth;
    unsigned status;               /* sticky global bits (overflow, internal error) */
    int sim_bits;                  /* key bits holding the simulation id */
    unsigned long long counters[5];/* body-body, body-cell, opened, wave node steps,
                                      group far-field (cell, group) pairs */
    int halt;                      /* a step ended with a fatal status: later steps do nothing */
    unsigned steps_done;           /* KDK steps completed on the device */
    int level_count[64];           /* cells per level */
    int level_fill[64];
    int level_off[65];             /* exclusive scan of level_count */
    unsigned group_mean;           /* group walk: mean predicted group cost of the last walk */
};

/* Walk tuning picked by the autotuner (nbhip_autotune). */
struct WalkTune {
    int block;                     /* threads per workgroup: 64, 128 or 256 */
    int batch;                     /* mixed walk: node block per FP32 partial sum (power of 2);
                                      group walk: LDS interaction-list depth (64, 128, 256) */
};

/* Group walk tasks: the sub-groups of a 64-body group predicted to cost
 * more than NBH_GW_SPLIT_COST times the mean group are shared out among up
 * to 8 waves (runs of whole sub-groups, each run at most that cost where a
 * sub-group allows); at most NBH_GW_EXTRA_TASKS such extra waves per walk. */
#define NBH_GW_EXTRA_TASKS 1024
#define NBH_GW_SPLIT_COST 2
__host__ __device__ inline int nbh_group_tasks(int n) { return (n + 63) / 64 + NBH_GW_EXTRA_TASKS; }

/* Min-hierarchy over the adjacent prefixes, in dadj's allocation after its
 * n entries: level 0 is dadj[0..M), level L >= 1 holds the minimum of each 16
 * entries of level L-1 (size ceil(M / 16^L)), every level starting on a
 * multiple of 16 entries, up to the first level of at most 16 entries. */
struct NbhDHier {
    int top;
    int size[9];
    int off[9];
};
__host__ __device__ inline NbhDHier nbh_dhier(int n, int M)
{
    NbhDHier h;
    h.top = 0;
    h.size[0] = M;
    h.off[0] = 0;
    int at = (n + 15) & ~15;
    while (h.size[h.top] > 16 && h.top < 8) {
        const int s = (h.size[h.top] + 15) / 16;
        ++h.top;
        h.size[h.top] = s;
        h.off[h.top] = at;
        at += (s + 15) & ~15;
    }
    return h;
}
/* dadj's allocation: n prefixes and the hierarchy's levels for M <= n */
__host__ __device__ inline size_t nbh_dadj_cap(int n)
{
    size_t at = ((size_t) n + 15) & ~(size_t) 15, s = (size_t) n;
    while (s > 16) {
        s = (s + 15) / 16;
        at += (s + 15) & ~(size_t) 15;
    }
    return at + 16;
}

#define NBH_SORT_TILE 2048   /* keys per radix-sort block: 256 threads x 8 */
#define NBH_SCAN_TILE 4096   /* items per scan block: 1024 threads x 4 */

/* Adjoint bookkeeping (FP64 states only): checkpoints of the forward run
 * every `every` steps, one recomputed segment, and the adjoint vectors.
 * A state slot holds 9 arrays of n doubles: x y z vx vy vz ax ay az. */
struct NbhAdjoint {
    unsigned nsteps, every, nckpt;
    double* ckpt;                 /* nckpt slots */
    SimTab* ckpt_tab;             /* nckpt * nsim */
    double* seg;                  /* every + 1 slots */
    SimTab* seg_tab;              /* (every + 1) * nsim */
    double* fin;                  /* final forward state, one slot */
    SimTab* fin_tab;
    double* g;                    /* gx gy gz | gvx gvy gvz | gax gay gaz | gm: 10 n */
    double* gdt;                  /* nsim */
    double* gnode;                /* 10 * node_cap: d/d(cm x y z, mass, quad[6]) */
    int ready;
};

struct nbhip_state {
    nbhip_ctx ctx;
    int device;
    int n;                        /* bodies, all simulations */
    int nsim;
    int sim_shift;                /* key bit where the simulation id starts */
    int top_levels;               /* Morton levels in the sort word */
    int node_cap;
    int sort_blocks;
    hipStream_t stream;

    double *px, *py, *pz, *vx, *vy, *vz, *m, *ax, *ay, *az;
    int *simid;                   /* simulation of each body, original order */

    unsigned long long *key, *key_alt;
    unsigned *perm, *perm_alt;
    unsigned *hist;
    unsigned *dtot;               /* per-digit totals of one radix pass */

    double4 *pos4;                /* x, y, z, m per body, original order (build input) */
    double4 *spos;
    int *simid_s;                 /* simulation of each sorted body */
    unsigned long long *su;       /* 3 per sorted body: ux, uy, uz */
    int *dadj;
    unsigned *cincl;              /* inclusive scan of cells-per-first-body */
    unsigned *scan_part;


    NodeW *node;
    double *nquad;                /* 6 per node */
    double *nmid;                 /* 3 per node (cells) */
    int *nnext;                   /* cells: pre-order index of Next(cell) (the walk record's
                                     next, kept compact for the summaries) */
    int *nlevel, *nparent, *nsim_node;
    int *level_list;              /* cell pre-order indices grouped by level */
    NodeF *nodef;                 /* mixed-precision walk records (precision MIXED) */
    NodeG *nodeg;                 /* group walk table (precision MIXED_GROUP): nsim + node_cap */
    unsigned *bcost;              /* group walk: per body (original order), its sub-group's work
                                     in the last walk */
    float *bbox;                  /* group walk: per body, its sub-group's box half-diagonal in
                                     the last walk */
    unsigned *gcost;              /* group walk: predicted work of each 64-body group, and
                                     its task count - 1 in bits 28..30 */
    uint2 *gtask;                 /* group walk: up to 8 tasks per group: (cost, lo | hi << 8) */
    int2 *gorder;                 /* group walk: [0].x = task count, then the tasks (group,
                                     lo | hi << 8) in dispatch order, costliest first */
    unsigned *nkid;               /* children per node */
    unsigned *kbase;              /* exclusive scan of nkid: each cell's child block */
    int *nrank;                   /* a node's rank among its parent's children (arrival order) */
    int *kid;                     /* child lists: kid[kbase[p] .. + nkid[p]) */
    int *blist;                   /* summaries: inside cells of each body block, by level */
    int *span, *span_sorted;      /* summaries: cells spanning body blocks / super-blocks */
    double *exact_part;           /* Exact criterion: per-segment partial sums */
    int exact_seg;                /* Exact criterion: most segments of any simulation */

    DevScalars *scal;
    SimTab *sims;
    int *h_body_off, *h_body_n;   /* host copies */
    int all_eps2_pos;

    unsigned step;
    unsigned dev_steps;           /* DevScalars::steps_done at the last status read */
    double* best_buf;             /* bodies of the best-likelihood step: x y z vx vy vz */
    int best_kept;                /* best_buf holds the last run's best step */
    unsigned status;              /* host-side sticky status (all simulations) */
    unsigned *h_sim_status;       /* host-side sticky status per simulation */
    int have_tree;
    int profiling;
    hipEvent_t ev[6];
    float phase_ms[5];

    int h_level_count[64];        /* cells per level of the last eager build */
    int h_levels_valid;
    int *h_counts_pinned;         /* pinned landing buffer of the level-count readback */
    hipEvent_t ev_counts;         /* recorded after that readback */

    NbhAdjoint* adj;

    unsigned char* ignore;        /* per body, original order: left out of histograms */
    unsigned* hist_buf;           /* device counts + totals of nbhip_histogram */
    size_t hist_cap;              /* unsigned entries in hist_buf */
    double open_scale;            /* coherent walk: critical radius^2 factor (>= 1) */
    double group_far;             /* group walk: far-field expansion ratio (0: off) */
    char* disp_buf;               /* scratch of nbhip_histogram_dispersion (grow-only) */
    size_t disp_cap;              /* bytes in disp_buf */

    WalkTune tune;
    int use_graph;
    hipGraph_t graph;
    hipGraphExec_t graph_exec;
    int graph_ready;
};

/* launchers (tree.hip) */
void nbh_bbox(nbhip_state* st, int drift);
void nbh_build(nbhip_state* st, int graph_safe);
unsigned nbh_level_lists(nbhip_state* st);
void nbh_step_end(nbhip_state* st);
/* launchers (walk.hip) */
void nbh_walk(nbhip_state* st, int kick, int count, double* out_ax, double* out_ay,
              double* out_az, int self_only, const WalkTune* tune);
void nbh_direct(nbhip_state* st, int nsample, const int* d_idx, double* d_acc);
/* launcher (walk_group.hip) */
struct WalkOut;
void nbh_walk_group(nbhip_state* st, const WalkOut& o, double* out_ax, double* out_ay,
                    double* out_az, const WalkTune& t, int count);
/* launcher (histogram.hip) */
void nbh_histogram(nbhip_state* st, const nbhip_histogram_params* hp, unsigned* d_counts,
                   unsigned* d_totals);
void nbh_dispersion(nbhip_state* st, int off, int cnt, const nbhip_histogram_params* hp,
                    const nbhip_disp_params* dp, int* dbin, double* dval, double* part,
                    double* state, double* out);
/* scratch of the allocation-free entry point (histogram.hip) */
unsigned nbh_scratch_init(int device);
/* launchers (adjoint.hip) */
void nbh_adj_kick_end(nbhip_state* st, const double* a_next);
void nbh_adj_force(nbhip_state* st);
void nbh_exact_force(nbhip_state* st, int kick, double* oax, double* oay, double* oaz,
                     int use_pot);
void nbh_exact_adjoint(nbhip_state* st, double* g);
void nbh_adj_body(nbhip_state* st, int drift_kick, const double* v_prev, const double* a_prev);

#define NBH_CHECK(x)                                                         \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) {                                              \
            nbh_set_error("%s:%d: %s: %s", __FILE__, __LINE__, #x,           \
                          hipGetErrorString(e_));                            \
            return NBHIP_GPU_ERROR;                                          \
        }                                                                    \
    } while (0)

void nbh_set_error(const char* fmt, ...);

#endif
