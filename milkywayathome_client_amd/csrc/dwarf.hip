/*
 * Two-component dwarf initial conditions on the GPU
 * (predefinedModels.mixeddwarf: nbGenerateMixedDwarfCore,
 * nbody/src/nbody_mixeddwarf.c:765-1026).
 *
 * The model is the reference's: the components' densities and potentials
 * (nbody/src/nbody_dwarf_potential.c: Plummer, NFW, general Hernquist),
 * set_vars / get_extra_nfw_mass (:696-762: an NFW mass is M200 and is reset
 * to the mass inside the sampling bound 5 r200; other components are bounded
 * at 50 (a_light + a_dark)), radii by the reference's rejection (r uniform on
 * [0, bound], accepted with probability r^2 rho(r) / max r^2 rho; r_mag,
 * :538-574), speeds by rejection against each component's Eddington
 * distribution function in the total potential (v uniform below 0.99 v_esc;
 * vel_mag, :576-618), isotropic directions by rejection from the cube
 * (get_components, :620-647; the velocity's first), and the per-component
 * centre-of-mass and momentum correction to the orbit point
 * (cm_correction_by_comp, :649-690).  Baryons come first.
 *
 * This project's own parts:
 *  - the distribution function is tabulated once per component by Eddington's
 *    formula, f(E) = 1/(sqrt(8) pi^2) int_0^E (d^2 rho / d psi^2) dpsi /
 *    sqrt(E - psi), with analytic derivatives of rho and psi in r and
 *    Gauss-Legendre quadrature after psi = E - t^2 (the reference evaluates
 *    a stencil-differenced integral, cut at 10 r(E), on every trial);
 *    the sampler interpolates log f linearly in log E;
 *  - every body draws from its own counter-based stream (Philox4x32-10,
 *    Salmon et al., SC'11, keyed by the seed, counter = body index and draw
 *    number), so the bodies are independent of each other, of the launch
 *    shape and of the thread count: one thread per body;
 *  - the centre-of-mass sums are taken in a fixed order (per block, then the
 *    blocks in order), so repeated runs are bit-identical.
 * The outputs are therefore statistically, not stream-, equal to the
 * reference's (whose sampler is serial dSFMT and takes hours at 10^6 bodies).
 */
#include "nbhip_internal.h"
#include "nbhip_rng.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr double kPcrit = 0.000679087369829744220469326744094105320596648627735869652;
constexpr double kVolPcrit = 0.568910904587397184785763397846734505212216314432372653620;
constexpr int kTable = 2048;        /* energies per distribution-function table */
constexpr int kGauss = 128;         /* Gauss-Legendre nodes of the Eddington integral */
constexpr int kScan = 128;          /* speeds scanned for the rejection envelope */
constexpr int kMaxTries = 1 << 22;  /* rejection trials per body before giving up */
constexpr int kSumBodies = 1024;    /* bodies per centre-of-mass partial sum */

struct Comp {
    int type;                       /* 0 Plummer, 1 NFW, 2 general Hernquist (dwarf_t) */
    double mass, a, p0;
};

struct Model {
    Comp c[2];                      /* light, dark */
    double bound[2];                /* sampling bounds */
    double r2rho_max[2];            /* max r^2 rho(r) of each component */
    double lnE0, dlnE;              /* the tables' energy grid: ln E_k = lnE0 + k dlnE */
    unsigned nbody, nbody_baryon;
    double pmass[2];                /* mass of one body of each component */
    unsigned long long seed;
};

/* density and its first two r-derivatives */
__host__ __device__ inline void nbh_rho(const Comp& c, double r, double* f, double* d1,
                                         double* d2)
{
    const double a = c.a;
    if (c.type == 0) {
        const double K = (3.0 / (4.0 * M_PI)) * (c.mass / (a * a * a));
        const double x = (r / a) * (r / a);
        const double s = 1.0 / (1.0 + x);
        const double s52 = s * s * sqrt(s);
        *f = K * s52;
        *d1 = -5.0 * K * (r / (a * a)) * s52 * s;
        *d2 = -5.0 * K / (a * a) * s52 * s * s * (1.0 - 6.0 * x);
    } else if (c.type == 1) {
        const double R = r / a;
        const double q = 1.0 + R;
        *f = c.p0 / (R * q * q);
        *d1 = -c.p0 * (1.0 + 3.0 * R) / (R * R * q * q * q) / a;
        *d2 = 2.0 * c.p0 * (6.0 * R * R + 4.0 * R + 1.0) / (R * R * R * q * q * q * q) / (a * a);
    } else {
        const double K = c.mass * a / (2.0 * M_PI);
        const double q = r + a;
        const double q4 = q * q * q * q;
        *f = K / (r * q * q * q);
        *d1 = -K * (4.0 * r + a) / (r * r * q4);
        *d2 = K * (20.0 * r * r + 10.0 * a * r + 2.0 * a * a) / (r * r * r * q4 * q);
    }
}

/* psi = -potential and its first two r-derivatives */
__host__ __device__ inline void nbh_psi(const Comp& c, double r, double* f, double* d1,
                                         double* d2)
{
    const double a = c.a;
    if (c.type == 0) {
        const double s = r * r + a * a;
        const double is = 1.0 / s;
        const double ir = sqrt(is);
        *f = c.mass * ir;
        *d1 = -c.mass * r * ir * is;
        *d2 = -c.mass * ir * is * is * (a * a - 2.0 * r * r);
    } else if (c.type == 1) {
        const double C = 4.0 * M_PI * a * a * c.p0;
        const double R = r / a;
        const double L = log1p(R);
        const double q = 1.0 + R;
        *f = C * L / R;
        *d1 = C * (1.0 / (R * q) - L / (R * R)) / a;
        *d2 = C * (-(1.0 + 2.0 * R) / (R * R * q * q) + 2.0 * L / (R * R * R) - 1.0 / (R * R * q)) /
              (a * a);
    } else {
        const double q = r + a;
        *f = c.mass / q;
        *d1 = -c.mass / (q * q);
        *d2 = 2.0 * c.mass / (q * q * q);
    }
}

__host__ __device__ inline double nbh_psi_total(const Model& m, double r)
{
    double f0, f1, d, e;
    nbh_psi(m.c[0], r, &f0, &d, &e);
    nbh_psi(m.c[1], r, &f1, &d, &e);
    return f0 + f1;
}

/* f_c(E) by Eddington's formula: 2 int_0^sqrt(E) g(E - t^2) dt / (sqrt(8) pi^2)
 * with g = d^2 rho_c / d psi^2 at the radius where the total psi equals
 * E - t^2 (found by bisection in ln r). */
__global__ void k_df_table(Model m, int comp, const double* __restrict__ gx,
                           const double* __restrict__ gw, double* __restrict__ lnf)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= kTable) return;
    const double E = exp(m.lnE0 + m.dlnE * k);
    const double se = sqrt(E);
    const double amin = fmin(m.c[0].a, m.c[1].a), amax = fmax(m.c[0].a, m.c[1].a);
    double acc = 0.0;
    for (int j = 0; j < kGauss; ++j) {
        const double t = 0.5 * se * (gx[j] + 1.0);
        const double psi = E - t * t;
        /* ln r with psi_total(r) = psi (psi decreases with r) */
        double lo = log(amin) - 30.0, hi = log(amax) + 40.0;
        for (int it = 0; it < 80; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (nbh_psi_total(m, exp(mid)) > psi) lo = mid; else hi = mid;
        }
        const double r = exp(0.5 * (lo + hi));
        double rf, r1, r2, p0, p1, p2, q0, q1, q2;
        nbh_rho(m.c[comp], r, &rf, &r1, &r2);
        nbh_psi(m.c[0], r, &p0, &p1, &p2);
        nbh_psi(m.c[1], r, &q0, &q1, &q2);
        const double d1 = p1 + q1, d2 = p2 + q2;
        const double g = (r2 * d1 - r1 * d2) / (d1 * d1 * d1);
        acc += gw[j] * g;
    }
    /* 2 * (se / 2) * sum = se * sum */
    const double f = se * acc / (sqrt(8.0) * M_PI * M_PI);
    lnf[k] = f > 0.0 ? log(f) : -INFINITY;
}

__device__ inline double nbh_df(const Model& m, const double* __restrict__ lnf, double E)
{
    if (!(E > 0.0)) return 0.0;
    double x = (log(E) - m.lnE0) / m.dlnE;
    x = fmin(fmax(x, 0.0), (double) (kTable - 1) - 1e-9);
    const int k = (int) x;
    const double w = x - k;
    const double a = lnf[k], b = lnf[k + 1];
    if (a == -INFINITY || b == -INFINITY) return 0.0;
    return exp(a + w * (b - a));
}

/* One body's stream: Philox4x32-10 blocks (counter = body, draw number),
 * two uniforms in (0, 1) per block. */
struct Stream {
    unsigned key[2];
    unsigned body;
    unsigned n;
    double buf[2];
    int left;
    __device__ double next()
    {
        if (left == 0) {
            const unsigned ctr[4] = { body, n++, 0u, 0x64776172u };
            unsigned o[4];
            nbh_philox4x32_10(ctr, key, o);
            buf[0] = nbh_u01(o[0], o[1]);
            buf[1] = nbh_u01(o[2], o[3]);
            left = 2;
        }
        return buf[2 - left--];
    }
};

/* rejection from the cube (mwRandomUnitPoint + the unit-sphere test) */
__device__ inline bool nbh_direction(Stream& s, double* u)
{
    for (int t = 0; t < kMaxTries; ++t) {
        const double x = 2.0 * s.next() - 1.0;
        const double y = 2.0 * s.next() - 1.0;
        const double z = 2.0 * s.next() - 1.0;
        const double r2 = x * x + y * y + z * z;
        if (r2 <= 1.0 && r2 > 0.0) {
            const double ir = 1.0 / sqrt(r2);
            u[0] = x * ir;
            u[1] = y * ir;
            u[2] = z * ir;
            return true;
        }
    }
    return false;
}

__global__ __launch_bounds__(256) void k_dwarf_sample(Model m, const double* __restrict__ lnf,
                                                      double* __restrict__ px,
                                                      double* __restrict__ py,
                                                      double* __restrict__ pz,
                                                      double* __restrict__ vx,
                                                      double* __restrict__ vy,
                                                      double* __restrict__ vz,
                                                      double* __restrict__ pm,
                                                      unsigned* __restrict__ fail)
{
    const unsigned i = blockIdx.x * 256u + threadIdx.x;
    if (i >= m.nbody) return;
    const int comp = i < m.nbody_baryon ? 0 : 1;
    const Comp c = m.c[comp];
    const double* tab = lnf + (size_t) comp * kTable;
    Stream s;
    s.key[0] = (unsigned) m.seed;
    s.key[1] = (unsigned) (m.seed >> 32);
    s.body = i;
    s.n = 0;
    s.left = 0;
    bool ok = false;
    /* radius: r_mag's rejection */
    double r = 0.0;
    for (int t = 0; t < kMaxTries && !ok; ++t) {
        r = s.next() * m.bound[comp];
        const double u = s.next();
        double f, d1, d2;
        nbh_rho(c, r, &f, &d1, &d2);
        ok = r > 0.0 && r * r * f / m.r2rho_max[comp] > u;
    }
    /* speed: rejection against v^2 f(psi(r) - v^2 / 2) below 0.99 v_esc,
     * under an envelope from a scan of the speeds */
    const double psi = nbh_psi_total(m, r);
    const double vesc = 0.99 * sqrt(fabs(2.0 * psi));
    double gmax = 0.0;
    for (int k = 1; k <= kScan; ++k) {
        const double v = vesc * k / kScan;
        gmax = fmax(gmax, v * v * nbh_df(m, tab, psi - 0.5 * v * v));
    }
    gmax *= 1.25;
    double v = 0.0;
    bool vok = false;
    for (int t = 0; t < kMaxTries && ok && !vok; ++t) {
        v = s.next() * vesc;
        const double u = s.next();
        const double g = v * v * nbh_df(m, tab, psi - 0.5 * v * v);
        vok = v > 0.0 && g > u * gmax;
        if (g > gmax) atomicOr(fail, 2u);            /* envelope too low */
    }
    double dv[3], dx[3];
    ok = ok && vok && nbh_direction(s, dv) && nbh_direction(s, dx);
    if (!ok) atomicOr(fail, 1u);
    px[i] = r * dx[0];
    py[i] = r * dx[1];
    pz[i] = r * dx[2];
    vx[i] = v * dv[0];
    vy[i] = v * dv[1];
    vz[i] = v * dv[2];
    pm[i] = m.pmass[comp];
}

/* Per-block sums of m x, m y, m z, m vx, m vy, m vz of each component (12
 * values), bodies in order within a thread, a fixed tree across threads. */
__global__ __launch_bounds__(256) void k_dwarf_sums(Model m, const double* __restrict__ px,
                                                    const double* __restrict__ py,
                                                    const double* __restrict__ pz,
                                                    const double* __restrict__ vx,
                                                    const double* __restrict__ vy,
                                                    const double* __restrict__ vz,
                                                    const double* __restrict__ pm,
                                                    double* __restrict__ part)
{
    __shared__ double sh[12][256];
    const int tid = threadIdx.x;
    double acc[12];
    for (int q = 0; q < 12; ++q) acc[q] = 0.0;
    const unsigned b0 = blockIdx.x * (unsigned) kSumBodies;
    for (int j = 0; j < kSumBodies / 256; ++j) {
        const unsigned i = b0 + (unsigned) j * 256u + (unsigned) tid;
        if (i >= m.nbody) break;
        const int o = i < m.nbody_baryon ? 0 : 6;
        const double w = pm[i];
        acc[o + 0] += w * px[i];
        acc[o + 1] += w * py[i];
        acc[o + 2] += w * pz[i];
        acc[o + 3] += w * vx[i];
        acc[o + 4] += w * vy[i];
        acc[o + 5] += w * vz[i];
    }
    for (int q = 0; q < 12; ++q) sh[q][tid] = acc[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h)
            for (int q = 0; q < 12; ++q) sh[q][tid] += sh[q][tid + h];
        __syncthreads();
    }
    if (tid < 12) part[12 * (size_t) blockIdx.x + tid] = sh[tid][0];
}

/* Blocks summed in order, divided by the component masses; then every body
 * moved (cm_correction_by_comp). */
__global__ void k_dwarf_total(int nblocks, const double* __restrict__ part,
                              double* __restrict__ cm, double m_light, double m_dark)
{
    const int q = threadIdx.x;
    if (q >= 12) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += part[12 * (size_t) b + q];
    const double mass = q < 6 ? m_light : m_dark;
    cm[q] = mass > 0.0 ? s / mass : 0.0;
}

__global__ __launch_bounds__(256) void k_dwarf_shift(Model m, const double* __restrict__ cm,
                                                     double rx, double ry, double rz, double ux,
                                                     double uy, double uz,
                                                     double* __restrict__ px,
                                                     double* __restrict__ py,
                                                     double* __restrict__ pz,
                                                     double* __restrict__ vx,
                                                     double* __restrict__ vy,
                                                     double* __restrict__ vz)
{
    const unsigned i = blockIdx.x * 256u + threadIdx.x;
    if (i >= m.nbody) return;
    const double* c = cm + (i < m.nbody_baryon ? 0 : 6);
    px[i] = px[i] - c[0] + rx;
    py[i] = py[i] - c[1] + ry;
    pz[i] = pz[i] - c[2] + rz;
    vx[i] = vx[i] - c[3] + ux;
    vy[i] = vy[i] - c[4] + uy;
    vz[i] = vz[i] - c[5] + uz;
}

/* Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_n). */
void nbh_gauss_legendre(int n, double* x, double* w)
{
    for (int i = 0; i < n; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                const double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            const double dz = p1 / pp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        x[i] = -z;
        w[i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
}

unsigned nbh_model(const nbhip_dwarf_params* p, Model* m)
{
    const int types[2] = { p->light_type, p->dark_type };
    const double mass[2] = { p->light_mass, p->dark_mass };
    const double scale[2] = { p->light_scale, p->dark_scale };
    if (p->nbody_baryon > p->nbody) {
        nbh_set_error("nbhip_mixed_dwarf: nbody_baryon %u > nbody %u", p->nbody_baryon, p->nbody);
        return NBHIP_USER_ERROR;
    }
    memset(m, 0, sizeof(*m));
    for (int k = 0; k < 2; ++k) {
        if (types[k] < NBHIP_DWARF_PLUMMER || types[k] > NBHIP_DWARF_GEN_HERNQUIST) {
            nbh_set_error("nbhip_mixed_dwarf: component type %d is not supported", types[k]);
            return NBHIP_UNSUPPORTED;
        }
        if (!(scale[k] > 0.0) || !(mass[k] >= 0.0)) {
            nbh_set_error("nbhip_mixed_dwarf: component %d needs a positive scale radius", k);
            return NBHIP_USER_ERROR;
        }
        Comp& c = m->c[k];
        c.type = types[k];
        c.mass = mass[k];
        c.a = scale[k];
        /* set_vars */
        const double r200 = cbrt(c.mass / kVolPcrit);
        const double cc = r200 / c.a;
        const double term = log(1.0 + cc) - cc / (1.0 + cc);
        c.p0 = 200.0 * cc * cc * cc * kPcrit / (3.0 * term);
        if (c.type == NBHIP_DWARF_NFW) {
            m->bound[k] = 5.0 * r200;
            const double b = m->bound[k];
            c.mass = 4.0 * M_PI * c.p0 * c.a * c.a * c.a * (log((c.a + b) / c.a) - b / (c.a + b));
        } else {
            m->bound[k] = 50.0 * (scale[0] + scale[1]);
        }
        /* where r^2 rho peaks (nbGenerateMixedDwarfCore's analytic maxima) */
        const double rpk = c.type == NBHIP_DWARF_PLUMMER ? sqrt(2.0 / 3.0) * c.a
                         : c.type == NBHIP_DWARF_NFW ? c.a : 0.5 * c.a;
        double f, d1, d2;
        nbh_rho(c, rpk, &f, &d1, &d2);
        m->r2rho_max[k] = rpk * rpk * f;
    }
    const unsigned nb = p->nbody_baryon, nd = p->nbody - p->nbody_baryon;
    m->pmass[0] = nb ? m->c[0].mass / nb : 0.0;
    m->pmass[1] = nd ? m->c[1].mass / nd : 0.0;
    m->nbody = p->nbody;
    m->nbody_baryon = nb;
    m->seed = p->seed;
    /* energies from 1 % of psi at the outer bound (the least a sampled body
     * can have: E >= (1 - 0.99^2) psi(r)) to psi at 1e-3 of the smaller
     * scale radius (above it the table's last value is used: closer in, a
     * Plummer-dominated centre makes rho'' psi' - rho' psi'' cancel to
     * leading order and the quotient loses eps (a / r)^2 relative; bodies
     * that far in are ~1e-7 of the dark matter and ~1e-9 of the baryons) */
    const double bmax = fmax(m->bound[0], m->bound[1]);
    const double amin = fmin(m->c[0].a, m->c[1].a);
    const double e_lo = 0.01 * nbh_psi_total(*m, bmax);
    const double e_hi = nbh_psi_total(*m, 1e-3 * amin);
    m->lnE0 = log(e_lo);
    m->dlnE = (log(e_hi) - m->lnE0) / (kTable - 1);
    return NBHIP_SUCCESS;
}

/* The tables of both components into d_lnf (2 x kTable). */
unsigned nbh_tables(const Model& m, double* d_lnf, double* d_gauss, hipStream_t s)
{
    std::vector<double> g(2 * kGauss);
    nbh_gauss_legendre(kGauss, g.data(), g.data() + kGauss);
    NBH_CHECK(hipMemcpyAsync(d_gauss, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice,
                             s));
    for (int comp = 0; comp < 2; ++comp)
        hipLaunchKernelGGL(k_df_table, dim3((kTable + 127) / 128), dim3(128), 0, s, m, comp,
                           d_gauss, d_gauss + kGauss, d_lnf + comp * kTable);
    NBH_CHECK(hipGetLastError());
    return NBHIP_SUCCESS;
}

/* Sample into device arrays (SoA); scratch holds the tables, Gauss nodes,
 * partial sums and flags. */
unsigned nbh_dwarf_device(const nbhip_dwarf_params* p, const Model& m, double* px, double* py,
                          double* pz, double* vx, double* vy, double* vz, double* pmass,
                          hipStream_t s)
{
    const unsigned n = m.nbody;
    const int nblk = (int) ((n + kSumBodies - 1) / kSumBodies);
    const size_t scratch_d = 2 * kTable + 2 * kGauss + 12 * (size_t) (nblk ? nblk : 1) + 12;
    double* scr = nullptr;
    unsigned* d_fail = nullptr;
    NBH_CHECK(hipMalloc((void**) &scr, scratch_d * sizeof(double)));
    hipError_t e = hipMalloc((void**) &d_fail, sizeof(unsigned));
    if (e != hipSuccess) {
        (void) hipFree(scr);
        NBH_CHECK(e);
    }
    double* lnf = scr;
    double* gauss = lnf + 2 * kTable;
    double* part = gauss + 2 * kGauss;
    double* cm = part + 12 * (size_t) (nblk ? nblk : 1);
    unsigned rc = nbh_tables(m, lnf, gauss, s);
    unsigned h_fail = 0;
    if (!rc && n) {
        e = hipMemsetAsync(d_fail, 0, sizeof(unsigned), s);
        if (e == hipSuccess) {
            const dim3 g((n + 255) / 256);
            hipLaunchKernelGGL(k_dwarf_sample, g, dim3(256), 0, s, m, lnf, px, py, pz, vx, vy, vz,
                               pmass, d_fail);
            hipLaunchKernelGGL(k_dwarf_sums, dim3(nblk), dim3(256), 0, s, m, px, py, pz, vx, vy,
                               vz, pmass, part);
            hipLaunchKernelGGL(k_dwarf_total, dim3(1), dim3(64), 0, s, nblk, part, cm,
                               m.c[0].mass, m.c[1].mass);
            hipLaunchKernelGGL(k_dwarf_shift, g, dim3(256), 0, s, m, cm, p->pos[0], p->pos[1],
                               p->pos[2], p->vel[0], p->vel[1], p->vel[2], px, py, pz, vx, vy,
                               vz);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(&h_fail, d_fail, sizeof(unsigned),
                                                hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            nbh_set_error("nbhip_mixed_dwarf: %s", hipGetErrorString(e));
            rc = NBHIP_GPU_ERROR;
        }
    }
    (void) hipFree(scr);
    (void) hipFree(d_fail);
    if (!rc && h_fail) {
        nbh_set_error("nbhip_mixed_dwarf: %s",
                      (h_fail & 1u) ? "a body was not accepted within the trial limit"
                                    : "a speed exceeded the rejection envelope");
        rc = NBHIP_ERROR;
    }
    return rc;
}

}  // namespace

extern "C" unsigned nbhip_philox4x32_10(const unsigned* ctr, const unsigned* key, unsigned* out)
{
    if (!ctr || !key || !out) return NBHIP_USER_ERROR;
    nbh_philox4x32_10(ctr, key, out);
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_mixed_dwarf(const nbhip_dwarf_params* p, int device, double* pos,
                                      double* vel, double* mass, double* comp_mass)
{
    if (!p || !pos || !vel || !mass) return NBHIP_USER_ERROR;
    Model m;
    unsigned rc = nbh_model(p, &m);
    if (rc) return rc;
    if (comp_mass) {
        comp_mass[0] = m.c[0].mass;
        comp_mass[1] = m.c[1].mass;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        nbh_set_error("nbhip_mixed_dwarf: no GPU %d", device);
        return NBHIP_CAPABILITY_ERROR;
    }
    NBH_CHECK(hipSetDevice(device));
    const size_t n = m.nbody;
    if (n == 0) return NBHIP_SUCCESS;
    double* d = nullptr;
    NBH_CHECK(hipMalloc((void**) &d, 7 * n * sizeof(double)));
    hipStream_t s;
    hipError_t e = hipStreamCreate(&s);
    if (e != hipSuccess) {
        (void) hipFree(d);
        NBH_CHECK(e);
    }
    rc = nbh_dwarf_device(p, m, d, d + n, d + 2 * n, d + 3 * n, d + 4 * n, d + 5 * n, d + 6 * n, s);
    std::vector<double> h;
    if (!rc) {
        h.resize(7 * n);
        e = hipMemcpyAsync(h.data(), d, 7 * n * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            nbh_set_error("nbhip_mixed_dwarf: %s", hipGetErrorString(e));
            rc = NBHIP_GPU_ERROR;
        }
    }
    (void) hipStreamDestroy(s);
    (void) hipFree(d);
    if (rc) return rc;
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            pos[3 * i + k] = h[k * n + i];
            vel[3 * i + k] = h[(3 + k) * n + i];
        }
        mass[i] = h[6 * n + i];
    }
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_mixed_dwarf_state(nbhip_state* st, int sim, const nbhip_dwarf_params* p,
                                            double* comp_mass)
{
    if (!st || !p || sim < 0 || sim >= st->nsim) return NBHIP_USER_ERROR;
    if ((int) p->nbody != st->h_body_n[sim]) {
        nbh_set_error("nbhip_mixed_dwarf_state: simulation %d holds %d bodies, not %u", sim,
                      st->h_body_n[sim], p->nbody);
        return NBHIP_USER_ERROR;
    }
    Model m;
    unsigned rc = nbh_model(p, &m);
    if (rc) return rc;
    if (comp_mass) {
        comp_mass[0] = m.c[0].mass;
        comp_mass[1] = m.c[1].mass;
    }
    NBH_CHECK(hipSetDevice(st->device));
    const size_t o = (size_t) st->h_body_off[sim];
    rc = nbh_dwarf_device(p, m, st->px + o, st->py + o, st->pz + o, st->vx + o, st->vy + o,
                          st->vz + o, st->m + o, st->stream);
    if (rc) return rc;
    /* the dark matter is left out of histograms (ignoreBody) */
    std::vector<unsigned char> ig;
    if (!st->ignore) {
        NBH_CHECK(hipMalloc((void**) &st->ignore, (size_t) (st->n ? st->n : 1)));
        NBH_CHECK(hipMemsetAsync(st->ignore, 0, (size_t) st->n, st->stream));
    }
    ig.assign(p->nbody, 0);
    for (unsigned i = p->nbody_baryon; i < p->nbody; ++i) ig[i] = 1;
    if (p->nbody)
        NBH_CHECK(hipMemcpyAsync(st->ignore + o, ig.data(), p->nbody, hipMemcpyHostToDevice,
                                 st->stream));
    NBH_CHECK(hipStreamSynchronize(st->stream));
    st->have_tree = 0;
    return NBHIP_SUCCESS;
}

extern "C" unsigned nbhip_dwarf_df(const nbhip_dwarf_params* p, int device, int comp, int count,
                                   const double* energy, double* f)
{
    if (!p || comp < 0 || comp > 1 || count < 0 || (count && (!energy || !f)))
        return NBHIP_USER_ERROR;
    Model m;
    unsigned rc = nbh_model(p, &m);
    if (rc) return rc;
    NBH_CHECK(hipSetDevice(device));
    double* d = nullptr;
    NBH_CHECK(hipMalloc((void**) &d, (2 * kTable + 2 * kGauss) * sizeof(double)));
    rc = nbh_tables(m, d, d + 2 * kTable, 0);
    std::vector<double> lnf(kTable);
    hipError_t e = hipSuccess;
    if (!rc) {
        e = hipMemcpy(lnf.data(), d + comp * kTable, kTable * sizeof(double),
                      hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            nbh_set_error("nbhip_dwarf_df: %s", hipGetErrorString(e));
            rc = NBHIP_GPU_ERROR;
        }
    }
    (void) hipFree(d);
    if (rc) return rc;
    for (int i = 0; i < count; ++i) {
        double x = (log(energy[i]) - m.lnE0) / m.dlnE;
        if (!(x >= 0.0) || x > kTable - 1) {
            f[i] = NAN;                     /* outside the table */
            continue;
        }
        const int k = x >= kTable - 1 ? kTable - 2 : (int) x;
        const double w = x - k;
        f[i] = exp(lnf[k] + w * (lnf[k + 1] - lnf[k]));
    }
    return NBHIP_SUCCESS;
}
