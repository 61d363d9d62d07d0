/*
 * Internal declarations of the nbhip library (not part of the C ABI).
 *
 * HBM layout of one simulation (N bodies, node capacity C):
 *   bodies, original order, SoA FP64 : px py pz vx vy vz m ax ay az   (80 B/body)
 *   sort: 64-bit Morton word (levels 1-21) + 32-bit body index, ping-pong
 *   sorted bodies: spos (x y z m, 32 B) and 63-bit fixed-point coordinates
 *     su (3 x 8 B) of the massive bodies, used for the 189-bit Morton prefix
 *   binary radix tree over the sorted massive bodies (Karras 2012):
 *     blo bhi bdelta (internal nodes), bparent (internal + leaves)
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
 * Position = root centre of mass + (hx + lx, hy + ly, hz + lz): the FP64
 * offset split into two floats, so that a float difference of two records
 * keeps ~1e-7 relative accuracy at any distance from the origin.  rcrit2 is
 * -1 for bodies (always accepted); q[6] is the quadrupole in FP32. */
struct __attribute__((aligned(16))) NodeF {
    float hx, hy, hz, m;
    float lx, ly, lz, rcrit2;
    int next, body;
    float q[6];
};

/* Device-resident scalars of one simulation. */
struct DevScalars {
    double rsize;                  /* st->tree.rsize, persists, grows only */
    double key_scale;              /* 2^63 / rsize */
    unsigned long long xyzmax_bits;/* max |x_i| of the last bbox pass */
    int nmassive;                  /* bodies with mass != 0 */
    int ncell;
    int nnode;
    int maxdepth;
    unsigned status;               /* sticky NBHIP_* bits of device checks */
    unsigned incest;               /* a body missed itself in the walk */
    unsigned long long counters[4];/* body-body, body-cell, opened, wave node steps */
    int level_count[64];           /* cells per level */
    int level_fill[64];
    int level_off[65];             /* exclusive scan of level_count */
};

#define NBH_SORT_TILE 4096   /* keys per radix-sort block: 256 threads x 16 */
#define NBH_SCAN_TILE 4096   /* items per scan block: 1024 threads x 4 */

struct nbhip_state {
    nbhip_ctx ctx;
    int device;
    int n;
    int node_cap;
    int sort_blocks;
    hipStream_t stream;

    double *px, *py, *pz, *vx, *vy, *vz, *m, *ax, *ay, *az;

    unsigned long long *key, *key_alt;
    unsigned *perm, *perm_alt;
    unsigned *hist;
    unsigned *dtot;               /* per-digit totals of one radix pass */

    double4 *spos;
    unsigned long long *su;       /* 3 * n: ux | uy | uz */
    int *dadj;
    unsigned *cincl;              /* inclusive scan of cells-per-first-body */
    unsigned *scan_part;

    int *blo, *bhi, *bdelta, *bparent;

    NodeW *node;
    double *nquad;                /* 6 per node */
    double *nmid;                 /* 3 per node (cells) */
    int *nlevel, *nparent;
    int *level_list;              /* cell pre-order indices grouped by level */
    NodeF *nodef;                 /* mixed-precision walk records (precision MIXED) */

    DevScalars *scal;

    unsigned step;
    unsigned status;              /* host-side sticky status */
    int have_tree;
    int profiling;
    hipEvent_t ev[6];
    float phase_ms[5];
};

/* launchers (tree.hip) */
void nbh_bbox(nbhip_state* st, int drift);
void nbh_build(nbhip_state* st);
/* launchers (walk.hip) */
void nbh_walk(nbhip_state* st, int kick, int count, double* out_ax, double* out_ay,
              double* out_az, int self_only);
void nbh_direct(nbhip_state* st, int nsample, const int* d_idx, double* d_acc);

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
