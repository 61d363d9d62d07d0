/*
 * CPU oracle: plain-C restatement of the reference nbody CPU path.
 * TEST INFRASTRUCTURE ONLY -- see nbody_oracle.h for what pins each part.
 *
 * Floating-point expressions keep the reference's operation order so that
 * results can be compared bit-for-bit; build with -ffp-contract=off.
 */
#include "nbody_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REF_NULL (-1)
#define CELL_REF(c) (-(c) - 2)
#define IS_CELL(r) ((r) <= -2)
#define IS_BODY(r) ((r) >= 0)
#define CELL_IDX(r) (-(r) - 2)

struct orc_tree {
    orc_cell* cells;
    int ncell;
    int cap;
    int* body_next;      /* Next() link of each body */
    int nbody_cap;
    double rsize;
    int max_depth;
    int structure_error;
    const double* bpos;  /* bodies of the last build */
    const double* bmass;
};

/* ------------------------------------------------------------------------ */
/* Correctly rounded helpers for the mw_pow calls of the potentials.       */
/* The reference maps mw_pow to crlibm pow_rn (milkyway_math_double.h:59):  */
/*   pow(x, 2.0) == x*x and pow(x, 0.5) == sqrt(x) exactly under correct    */
/*   rounding; x^1.5 and x^3 are formed in double-double and rounded once.  */
/* ------------------------------------------------------------------------ */
static double orc_pow1p5(double x)
{
    double s = sqrt(x);
    double e = fma(-s, s, x);          /* exact x - s*s */
    double ds = e / (2.0 * s);
    double p = x * s;
    double pe = fma(x, s, -p);         /* exact low word of x*s */
    return p + (pe + x * ds);
}

static double orc_cube(double t)
{
    double h = t * t;
    double l = fma(t, t, -h);
    double p = h * t;
    double pe = fma(h, t, -p);
    return p + (pe + l * t);
}

/* ------------------------------------------------------------------------ */
/* Tree construction: nbody/src/nbody_tree.c                               */
/* ------------------------------------------------------------------------ */
orc_tree* orc_tree_new(void)
{
    orc_tree* t = (orc_tree*) calloc(1, sizeof(orc_tree));
    return t;
}

void orc_tree_free(orc_tree* t)
{
    if (!t) return;
    free(t->cells);
    free(t->body_next);
    free(t);
}

/* nbMakeCell (nbody_tree.c:179-197) */
static int orc_make_cell(orc_tree* t)
{
    if (t->ncell == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 1024;
        t->cells = (orc_cell*) realloc(t->cells, (size_t) t->cap * sizeof(orc_cell));
    }
    orc_cell* c = &t->cells[t->ncell];
    memset(c, 0, sizeof(*c));
    for (int i = 0; i < 8; ++i) c->sub[i] = REF_NULL;
    c->more = REF_NULL;
    c->next = REF_NULL;
    return t->ncell++;
}

/* nbSubIndex (nbody_tree.c:36-52): octant bit set when midpoint <= body. */
static inline int orc_sub_index(const double* p, const double* q)
{
    int ind = 0;
    if (q[0] <= p[0]) ind += 4;
    if (q[1] <= p[1]) ind += 2;
    if (q[2] <= p[2]) ind += 1;
    return ind;
}

/* calcOffset (nbody_tree.c:226-231) */
static inline double orc_offset(double p, double q, double qsize)
{
    return q + 0.25 * (p < q ? -qsize : qsize);
}

/* expandBox (nbody_tree.c:156-176): root stays at the origin. */
static double orc_expand_box(double rsize, int n, const double* pos)
{
    double xyzmax = 0.0;
    for (int i = 0; i < n; ++i) {
        xyzmax = fmax(xyzmax, fabs(pos[3 * i + 0] - 0.0));
        xyzmax = fmax(xyzmax, fabs(pos[3 * i + 1] - 0.0));
        xyzmax = fmax(xyzmax, fabs(pos[3 * i + 2] - 0.0));
    }
    while (rsize < 2.0 * xyzmax) rsize *= 2.0;
    return rsize;
}

static inline const double* orc_ref_pos(const orc_tree* t, int r)
{
    return IS_CELL(r) ? t->cells[CELL_IDX(r)].pos : &t->bpos[3 * r];
}

static inline double orc_ref_mass(const orc_tree* t, int r)
{
    return IS_CELL(r) ? t->cells[CELL_IDX(r)].mass : t->bmass[r];
}

/* nbLoadBody (nbody_tree.c:242-282) */
static void orc_load_body(orc_tree* t, int b)
{
    const double* p = &t->bpos[3 * b];
    int q = 0;
    int qind = orc_sub_index(p, t->cells[q].pos);
    double qsize = t->rsize;
    int lev = 0;

    while (t->cells[q].sub[qind] != REF_NULL) {
        if (qsize <= DBL_EPSILON) {
            t->structure_error = 1;
            return;
        }
        int occ = t->cells[q].sub[qind];
        if (IS_BODY(occ)) {
            int c = orc_make_cell(t);
            orc_cell* cc = &t->cells[c];
            const double* qp = t->cells[q].pos;
            cc->pos[0] = orc_offset(p[0], qp[0], qsize);
            cc->pos[1] = orc_offset(p[1], qp[1], qsize);
            cc->pos[2] = orc_offset(p[2], qp[2], qsize);
            cc->level = lev + 1;
            cc->sub[orc_sub_index(&t->bpos[3 * occ], cc->pos)] = occ;
            t->cells[q].sub[qind] = CELL_REF(c);
        }
        q = CELL_IDX(t->cells[q].sub[qind]);
        qind = orc_sub_index(p, t->cells[q].pos);
        qsize *= 0.5;
        ++lev;
    }
    t->cells[q].sub[qind] = b;
    if (lev > t->max_depth) t->max_depth = lev;
}

/* bmax2Inc (nbody_tree.c:284-290) */
static inline double orc_bmax2_inc(double cm, double p, double psize)
{
    double dmin = cm - (p - 0.5 * psize);
    double m = fmax(dmin, psize - dmin);
    return m * m;
}

/* findRCrit (nbody_tree.c:308-343) */
static double orc_rcrit2(const orc_ctx* ctx, const double* mid, double rsize,
                         const double* cm, double psize)
{
    double rc;
    if (ctx->theta == 0.0) {
        rc = 2.0 * rsize;
        return rc * rc;
    }
    switch (ctx->criterion) {
    case ORC_TREECODE: {
        double dx = cm[0] - mid[0], dy = cm[1] - mid[1], dz = cm[2] - mid[2];
        rc = psize / ctx->theta + sqrt(dx * dx + dy * dy + dz * dz);
        return rc * rc;
    }
    case ORC_SW93: {
        double bmax2 = orc_bmax2_inc(cm[0], mid[0], psize);
        bmax2 += orc_bmax2_inc(cm[1], mid[1], psize);
        bmax2 += orc_bmax2_inc(cm[2], mid[2], psize);
        return bmax2 / (ctx->theta * ctx->theta);
    }
    case ORC_BH86:
    default:
        rc = psize / ctx->theta;
        return rc * rc;
    }
}

/* hackCofM + nbCheckTreeStructure (nbody_tree.c:345-424) */
static void orc_cofm(const orc_ctx* ctx, orc_tree* t, int p, double psize)
{
    double cm[3] = { 0.0, 0.0, 0.0 };
    double mass = 0.0;

    for (int i = 0; i < 8; ++i) {
        int q = t->cells[p].sub[i];
        if (q == REF_NULL) continue;
        if (IS_CELL(q)) orc_cofm(ctx, t, CELL_IDX(q), 0.5 * psize);
        double mq = orc_ref_mass(t, q);
        const double* pq = orc_ref_pos(t, q);
        mass += mq;
        cm[0] += mq * pq[0];
        cm[1] += mq * pq[1];
        cm[2] += mq * pq[2];
    }

    orc_cell* c = &t->cells[p];
    c->mass = mass;
    if (mass > 0.0) {
        cm[0] /= mass;
        cm[1] /= mass;
        cm[2] /= mass;
    } else {
        cm[0] = c->pos[0];
        cm[1] = c->pos[1];
        cm[2] = c->pos[2];
    }

    double half = 0.5 * psize;
    for (int d = 0; d < 3; ++d) {
        if (cm[d] < c->pos[d] - half || cm[d] > c->pos[d] + half)
            t->structure_error = 1;
    }

    memcpy(c->mid, c->pos, sizeof(c->mid));
    c->rcrit2 = orc_rcrit2(ctx, c->pos, t->rsize, cm, psize);
    memcpy(c->pos, cm, sizeof(cm));
}

/* threadTree (nbody_tree.c:127-150) */
static void orc_thread(orc_tree* t, int p, int n)
{
    if (IS_BODY(p)) {
        t->body_next[p] = n;
        return;
    }
    orc_cell* c = &t->cells[CELL_IDX(p)];
    c->next = n;
    int desc[9];
    int nd = 0;
    for (int i = 0; i < 8; ++i)
        if (c->sub[i] != REF_NULL) desc[nd++] = c->sub[i];
    c->more = nd ? desc[0] : n;
    desc[nd] = n;
    for (int i = 0; i < nd; ++i) orc_thread(t, desc[i], desc[i + 1]);
}

/* hackQuad (nbody_tree.c:70-121).  The moment accumulates from zero. */
static void orc_quad(orc_tree* t, int p)
{
    int desc[8];
    int nd = 0;
    for (int i = 0; i < 8; ++i)
        if (t->cells[p].sub[i] != REF_NULL) desc[nd++] = t->cells[p].sub[i];

    double Q[6] = { 0, 0, 0, 0, 0, 0 };
    for (int i = 0; i < nd; ++i) {
        int q = desc[i];
        if (IS_CELL(q)) orc_quad(t, CELL_IDX(q));
        const double* pq = orc_ref_pos(t, q);
        const double* pp = t->cells[p].pos;
        double dx = pq[0] - pp[0], dy = pq[1] - pp[1], dz = pq[2] - pp[2];
        double drsq = dz * dz + (dy * dy + dx * dx);
        double m = orc_ref_mass(t, q);
        double qd[6];
        qd[0] = m * (3.0 * (dx * dx) - drsq);
        qd[1] = m * (3.0 * (dx * dy));
        qd[2] = m * (3.0 * (dx * dz));
        qd[3] = m * (3.0 * (dy * dy) - drsq);
        qd[4] = m * (3.0 * (dy * dz));
        qd[5] = m * (3.0 * (dz * dz) - drsq);
        if (IS_CELL(q)) {
            const double* qq = t->cells[CELL_IDX(q)].quad;
            for (int k = 0; k < 6; ++k) qd[k] += qq[k];
        }
        for (int k = 0; k < 6; ++k) Q[k] += qd[k];
    }
    memcpy(t->cells[p].quad, Q, sizeof(Q));
}

/* nbMakeTree (nbody_tree.c:429-460) */
int orc_build_tree(const orc_ctx* ctx, int n, const double* pos, const double* mass,
                   double* rsize_io, orc_tree* t)
{
    t->ncell = 0;
    t->max_depth = 0;
    t->structure_error = 0;
    t->bpos = pos;
    t->bmass = mass;
    if (t->nbody_cap < n) {
        t->nbody_cap = n;
        t->body_next = (int*) realloc(t->body_next, (size_t) (n > 0 ? n : 1) * sizeof(int));
    }
    for (int i = 0; i < n; ++i) t->body_next[i] = REF_NULL;

    int root = orc_make_cell(t);
    t->cells[root].pos[0] = t->cells[root].pos[1] = t->cells[root].pos[2] = 0.0;
    t->cells[root].level = 0;

    *rsize_io = orc_expand_box(*rsize_io, n, pos);
    t->rsize = *rsize_io;

    for (int i = 0; i < n; ++i) {
        if (mass[i] != 0.0) orc_load_body(t, i);
        if (t->structure_error) return ORC_TREE_STRUCTURE_ERROR;
    }

    orc_cofm(ctx, t, root, t->rsize);
    if (t->structure_error) return ORC_TREE_STRUCTURE_ERROR;

    orc_thread(t, CELL_REF(root), REF_NULL);
    if (ctx->use_quad) orc_quad(t, root);
    return ORC_SUCCESS;
}

int orc_tree_ncell(const orc_tree* t) { return t->ncell; }
int orc_tree_maxdepth(const orc_tree* t) { return t->max_depth; }

static int orc_ref_next(const orc_tree* t, int r)
{
    return IS_CELL(r) ? t->cells[CELL_IDX(r)].next : t->body_next[r];
}

int orc_tree_nnode(const orc_tree* t)
{
    int cnt = 0;
    int q = CELL_REF(0);
    while (q != REF_NULL) {
        ++cnt;
        if (IS_CELL(q) && t->cells[CELL_IDX(q)].more != t->cells[CELL_IDX(q)].next)
            q = t->cells[CELL_IDX(q)].more;
        else
            q = orc_ref_next(t, q);
    }
    return cnt;
}

void orc_tree_preorder(const orc_tree* t, const double* bpos, const double* bmass,
                       int* kind, int* ref, int* skip, int* level,
                       double* pos, double* mass, double* rcrit2, double* quad, double* mid)
{
    int nn = orc_tree_nnode(t);
    int nb = t->nbody_cap;
    int* cell_pre = (int*) malloc((size_t) (t->ncell + 1) * sizeof(int));
    int* body_pre = (int*) malloc((size_t) (nb + 1) * sizeof(int));
    int* order = (int*) malloc((size_t) (nn + 1) * sizeof(int));
    int i = 0;
    int q = CELL_REF(0);
    while (q != REF_NULL) {
        order[i] = q;
        if (IS_CELL(q)) cell_pre[CELL_IDX(q)] = i; else body_pre[q] = i;
        ++i;
        if (IS_CELL(q) && t->cells[CELL_IDX(q)].more != t->cells[CELL_IDX(q)].next)
            q = t->cells[CELL_IDX(q)].more;
        else
            q = orc_ref_next(t, q);
    }
    for (i = 0; i < nn; ++i) {
        int r = order[i];
        int nx = orc_ref_next(t, r);
        skip[i] = nx == REF_NULL ? nn : (IS_CELL(nx) ? cell_pre[CELL_IDX(nx)] : body_pre[nx]);
        if (IS_CELL(r)) {
            const orc_cell* c = &t->cells[CELL_IDX(r)];
            kind[i] = 0;
            ref[i] = CELL_IDX(r);
            level[i] = c->level;
            memcpy(&pos[3 * i], c->pos, 3 * sizeof(double));
            memcpy(&mid[3 * i], c->mid, 3 * sizeof(double));
            mass[i] = c->mass;
            rcrit2[i] = c->rcrit2;
            memcpy(&quad[6 * i], c->quad, 6 * sizeof(double));
        } else {
            kind[i] = 1;
            ref[i] = r;
            level[i] = -1;
            memcpy(&pos[3 * i], &bpos[3 * r], 3 * sizeof(double));
            memset(&mid[3 * i], 0, 3 * sizeof(double));
            mass[i] = bmass[r];
            rcrit2[i] = 0.0;
            memset(&quad[6 * i], 0, 6 * sizeof(double));
        }
    }
    free(cell_pre);
    free(body_pre);
    free(order);
}

/* ------------------------------------------------------------------------ */
/* External potential: nbody/src/nbody_potential.c                         */
/* ------------------------------------------------------------------------ */
void orc_ext_accel(const orc_potential* pot, const double* pos, double* acc)
{
    const double x = pos[0], y = pos[1], z = pos[2];
    const double limit = 0.00390625; /* mw_pow(2.0, -8.0) */
    double ab = sqrt(z * z + (y * y + x * x));
    double r = (double) (ab <= limit) * limit + (double) (ab > limit) * ab;  /* :630-633 */
    double a[3] = { 0.0, 0.0, 0.0 };
    double t[3];

    /* miyamotoNagaiDiskAccel (:173-194) for disk and disk2 */
    for (int which = 0; which < 2; ++which) {
        int type = which ? pot->disk2_type : pot->disk_type;
        double M = which ? pot->disk2_mass : pot->disk_mass;
        double da = which ? pot->disk2_a : pot->disk_a;
        double db = which ? pot->disk2_b : pot->disk_b;
        if (type == ORC_DISK_MN) {
            double zp = sqrt(z * z + db * db);
            double azp = da + zp;
            double rp = x * x + y * y + azp * azp;
            double rth = orc_pow1p5(rp);
            t[0] = -M * x / rth;
            t[1] = -M * y / rth;
            t[2] = -M * z * azp / (zp * rth);
        } else {
            t[0] = t[1] = t[2] = 0.0;
        }
        if (which == 0) { a[0] = t[0]; a[1] = t[1]; a[2] = t[2]; }
        else { a[0] += t[0]; a[1] += t[1]; a[2] += t[2]; }
    }

    /* logHaloAccel (:488-504), nfwHaloAccel (:506-515) */
    if (pot->halo_type == ORC_HALO_LOG) {
        double v0 = pot->halo_vhalo, q = pot->halo_q, d = pot->halo_d;
        double zq = z / q;
        double denom = d * d + x * x + y * y + zq * zq;
        double k = -2.0 * v0 * v0 / denom;
        t[0] = k * x;
        t[1] = k * y;
        t[2] = k * z / (q * q);
    } else if (pot->halo_type == ORC_HALO_NFW) {
        double ha = pot->halo_d;
        double M = pot->halo_vhalo * pot->halo_vhalo * ha / 0.2162165954;
        double ar = ha + r;
        double c = (-M / (r * r)) * (log(ar / ha) / r - 1 / ar);
        t[0] = c * x; t[1] = c * y; t[2] = c * z;
    } else {
        t[0] = t[1] = t[2] = 0.0;
    }
    a[0] += t[0]; a[1] += t[1]; a[2] += t[2];

    /* hernquistSphericalAccel (:157-162), plummerSphericalAccel (:164-169) */
    if (pot->sphere_type == ORC_SPH_HERNQUIST) {
        double tmp = pot->sphere_scale + r;
        double s = -pot->sphere_mass / (r * (tmp * tmp));
        t[0] = s * x; t[1] = s * y; t[2] = s * z;
    } else if (pot->sphere_type == ORC_SPH_PLUMMER) {
        double sc = pot->sphere_scale;
        double tmp = sqrt(sc * sc + r * r);
        double s = -pot->sphere_mass / orc_cube(tmp);
        t[0] = s * x; t[1] = s * y; t[2] = s * z;
    } else {
        t[0] = t[1] = t[2] = 0.0;
    }
    a[0] += t[0]; a[1] += t[1]; a[2] += t[2];

    acc[0] = a[0]; acc[1] = a[1]; acc[2] = a[2];
}

/* ------------------------------------------------------------------------ */
/* Walk: nbGravity (nbody/src/nbody_grav.c:42-124)                         */
/* ------------------------------------------------------------------------ */
static void orc_gravity(const orc_ctx* ctx, const orc_tree* t, int p, const double* pos0,
                        double* acc, int* met_self, long long* cnt)
{
    double ax = 0.0, ay = 0.0, az = 0.0;
    int skip_self = 0;
    int q = CELL_REF(0);
    while (q != REF_NULL) {
        const double* pq = orc_ref_pos(t, q);
        double dx = pq[0] - pos0[0];
        double dy = pq[1] - pos0[1];
        double dz = pq[2] - pos0[2];
        double drSq = dz * dz + (dy * dy + dx * dx);
        int is_cell = IS_CELL(q);
        if (!is_cell || drSq >= t->cells[CELL_IDX(q)].rcrit2) {
            if (q != p) {
                drSq += ctx->eps2;
                double drab = sqrt(drSq);
                double phii = orc_ref_mass(t, q) / drab;
                double mor3 = phii / drSq;
                ax += mor3 * dx;
                ay += mor3 * dy;
                az += mor3 * dz;
                if (ctx->use_quad && is_cell) {
                    const double* Q = t->cells[CELL_IDX(q)].quad;
                    double qx = Q[0] * dx + Q[1] * dy + Q[2] * dz;
                    double qy = Q[1] * dx + Q[3] * dy + Q[4] * dz;
                    double qz = Q[2] * dx + Q[4] * dy + Q[5] * dz;
                    double drQdr = qx * dx + qy * dy + qz * dz;
                    double dr5inv = 1.0 / ((drSq * drSq) * drab);
                    double phiQ = 2.5 * (dr5inv * drQdr) / drSq;
                    ax += phiQ * dx;
                    ay += phiQ * dy;
                    az += phiQ * dz;
                    ax -= dr5inv * qx;
                    ay -= dr5inv * qy;
                    az -= dr5inv * qz;
                }
                if (cnt) { if (is_cell) cnt[1]++; else cnt[0]++; }
            } else {
                skip_self = 1;
            }
            q = orc_ref_next(t, q);
        } else {
            if (cnt) cnt[2]++;
            q = t->cells[CELL_IDX(q)].more;
        }
    }
    acc[0] = ax; acc[1] = ay; acc[2] = az;
    *met_self = skip_self;
}

/* nbMapForceBody (nbody_grav.c:126-203) */
static void orc_map_force(const orc_ctx* ctx, orc_state* st)
{
    const int n = st->n;
    int incest = 0;
    #pragma omp parallel for schedule(dynamic, 128) reduction(|: incest)
    for (int i = 0; i < n; ++i) {
        double a[3];
        int met;
        orc_gravity(ctx, st->tree, i, &st->pos[3 * i], a, &met, NULL);
        if (!met) incest = 1;
        if (ctx->use_potential) {
            double e[3];
            orc_ext_accel(&ctx->pot, &st->pos[3 * i], e);
            a[0] += e[0]; a[1] += e[1]; a[2] += e[2];
        }
        st->acc[3 * i + 0] = a[0];
        st->acc[3 * i + 1] = a[1];
        st->acc[3 * i + 2] = a[2];
    }
    if (incest) st->tree_incest = 1;
}

/* nbGravity_Exact (nbody_grav.c:205-227) for the bodies in idx. */
void orc_direct_accel(int n, const double* pos, const double* mass, double eps2,
                      int nsample, const int* idx, double* acc)
{
    #pragma omp parallel for schedule(dynamic, 16)
    for (int s = 0; s < nsample; ++s) {
        const double* p = &pos[3 * idx[s]];
        double ax = 0.0, ay = 0.0, az = 0.0;
        for (int i = 0; i < n; ++i) {
            double dx = pos[3 * i + 0] - p[0];
            double dy = pos[3 * i + 1] - p[1];
            double dz = pos[3 * i + 2] - p[2];
            double drSq = dz * dz + (dy * dy + dx * dx) + eps2;
            double drab = sqrt(drSq);
            double phii = mass[i] / drab;
            double mor3 = phii / drSq;
            ax += mor3 * dx;
            ay += mor3 * dy;
            az += mor3 * dz;
        }
        acc[3 * s + 0] = ax;
        acc[3 * s + 1] = ay;
        acc[3 * s + 2] = az;
    }
}

/* nbMapForceBody_Exact (nbody_grav.c:229-293) */
static void orc_map_force_exact(const orc_ctx* ctx, orc_state* st)
{
    const int n = st->n;
    #pragma omp parallel for schedule(dynamic, 128)
    for (int i = 0; i < n; ++i) {
        double a[3];
        orc_direct_accel(n, st->pos, st->mass, ctx->eps2, 1, &i, a);
        if (ctx->use_potential) {
            double e[3];
            orc_ext_accel(&ctx->pot, &st->pos[3 * i], e);
            a[0] += e[0]; a[1] += e[1]; a[2] += e[2];
        }
        st->acc[3 * i + 0] = a[0];
        st->acc[3 * i + 1] = a[1];
        st->acc[3 * i + 2] = a[2];
    }
}

/* nbIncestStatusCheck + nbGravMap (nbody_grav.c:295-328) */
int orc_grav_map(const orc_ctx* ctx, orc_state* st)
{
    if (ctx->criterion != ORC_EXACT) {
        if (!st->tree) st->tree = orc_tree_new();
        int rc = orc_build_tree(ctx, st->n, st->pos, st->mass, &st->rsize, st->tree);
        if (rc != ORC_SUCCESS) return rc;
        orc_map_force(ctx, st);
    } else {
        orc_map_force_exact(ctx, st);
    }
    if (st->tree_incest)
        return ctx->allow_incest ? ORC_TREE_INCEST_NONFATAL : ORC_TREE_INCEST_FATAL;
    return ORC_SUCCESS;
}

void orc_walk_bodies(const orc_ctx* ctx, const orc_state* st, int nsample,
                     const int* idx, double* acc, int* incest)
{
    int inc = 0;
    #pragma omp parallel for schedule(dynamic, 16) reduction(|: inc)
    for (int s = 0; s < nsample; ++s) {
        int met;
        orc_gravity(ctx, st->tree, idx[s], &st->pos[3 * idx[s]], &acc[3 * s], &met, NULL);
        if (!met) inc = 1;
    }
    *incest = inc;
}

void orc_walk_counts(const orc_ctx* ctx, const orc_state* st, int nsample,
                     const int* idx, long long* counts3)
{
    long long c0 = 0, c1 = 0, c2 = 0;
    #pragma omp parallel for schedule(dynamic, 16) reduction(+: c0, c1, c2)
    for (int s = 0; s < nsample; ++s) {
        long long c[3] = { 0, 0, 0 };
        double a[3];
        int met;
        orc_gravity(ctx, st->tree, idx[s], &st->pos[3 * idx[s]], a, &met, c);
        c0 += c[0]; c1 += c[1]; c2 += c[2];
    }
    counts3[0] = c0; counts3[1] = c1; counts3[2] = c2;
}

/* ------------------------------------------------------------------------ */
/* Integrator: nbody/src/nbody_plain.c:251-364                              */
/* ------------------------------------------------------------------------ */
orc_state* orc_state_new(int n, const double* pos, const double* vel, const double* mass,
                         double tree_rsize)
{
    orc_state* st = (orc_state*) calloc(1, sizeof(orc_state));
    size_t nn = (size_t) (n > 0 ? n : 1);
    st->n = n;
    st->pos = (double*) malloc(3 * nn * sizeof(double));
    st->vel = (double*) malloc(3 * nn * sizeof(double));
    st->mass = (double*) malloc(nn * sizeof(double));
    st->acc = (double*) calloc(3 * nn, sizeof(double));
    memcpy(st->pos, pos, 3 * (size_t) n * sizeof(double));
    memcpy(st->vel, vel, 3 * (size_t) n * sizeof(double));
    memcpy(st->mass, mass, (size_t) n * sizeof(double));
    st->rsize = tree_rsize;
    return st;
}

void orc_state_free(orc_state* st)
{
    if (!st) return;
    free(st->pos); free(st->vel); free(st->mass); free(st->acc);
    orc_tree_free(st->tree);
    free(st);
}

/* nbStepSystemPlain with zero LMC shifts (nbody_plain.c:329-364) */
int orc_step(const orc_ctx* ctx, orc_state* st)
{
    const double dt = ctx->timestep;
    const double dtHalf = 0.5 * dt;
    const int n = st->n;

    /* advancePosVel (:268-284) */
    #pragma omp parallel for schedule(dynamic, 128)
    for (int i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            double a = st->acc[3 * i + d] + 0.0;
            st->vel[3 * i + d] += dtHalf * a;
        }
        for (int d = 0; d < 3; ++d) st->pos[3 * i + d] += dt * st->vel[3 * i + d];
    }

    int rc = orc_grav_map(ctx, st);

    /* advanceVelocities (:301-315) */
    #pragma omp parallel for schedule(dynamic, 128)
    for (int i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            double a = st->acc[3 * i + d] + 0.0;
            st->vel[3 * i + d] += dtHalf * a;
        }
    }
    st->step++;
    return rc;
}

static int orc_fatal(int rc) { return (rc & ~(int) 0xFF000000u) != 0; }

/* nbRunSystemPlain (nbody_plain.c:366-461) without output hooks:
 * one force evaluation before the loop, then nsteps KDK steps. */
int orc_run(const orc_ctx* ctx, orc_state* st, unsigned nsteps)
{
    int rc = orc_grav_map(ctx, st);
    if (orc_fatal(rc)) return rc;
    for (unsigned k = 0; k < nsteps; ++k) {
        rc |= orc_step(ctx, st);
        if (orc_fatal(rc)) return rc;
    }
    return rc;
}

void orc_set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void) nthreads;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
