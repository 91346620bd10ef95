/*
 * cosenu_oracle.c -- CPU restatement of COSEnu's time-step path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP library is tested against this file:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  The product path never calls it.
 *
 * It restates, in the reference's floating-point operation order, with no
 * contraction (built with -ffp-contract=off), the following reference code:
 *   grid set-up            src/nuosc.hpp:78-96
 *   ghost fill             src/nuosc.hpp:212-278   (periodic and open)
 *   RK4 stage algebra      src/nuosc.hpp:280-348
 *   WENO7 reconstruction   src/rhs_fv.hpp:6-95
 *   FV right-hand side     src/rhs_fv.hpp:115-236
 *   FD right-hand side     src/rhs_fd.hpp:6-148    (ADVEC_CENTER_FD branch)
 *   initial condition      src/initialize.hpp:6-47, src/miscellaneous_funcs.hpp:57-77
 *   diagnostics            src/analysis.hpp:27-185
 * plus a moment-form variant of the collective integral (sum_k (1-v v_k) f_k
 * = F0 - v F1), used as an independent algebraic cross-check.
 *
 * Pinning (see tests/test_oracle_*.py): every function here is compared
 * bitwise with a build of the reference's own headers (oracle/_ref) on small
 * grids, and against closed forms (constant-state WENO exactness, 7th-order
 * convergence, homogeneous vacuum oscillation, rho_ex = 0 fixed point, trace
 * and lepton-number conservation, moment form = brute force).
 *
 * State exchange format: state[f][i][j], f in {ee, xx, ex_re, ex_im, bee,
 * bxx, bex_re, bex_im}, i = velocity bin (0..nvz-1), j = z cell (0..nz-1).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { ORC_FV = 0, ORC_FD = 1 };
enum { ORC_PERIODIC = 0, ORC_OPEN = 1 };
enum {
    ORC_VAC = 1,       /* VAC_OSC_ON */
    ORC_MAT = 2,       /* MAT_OSC_ON */
    ORC_COLL = 4,      /* COLL_OSC_ON */
    ORC_ADVEC_OFF = 8, /* ADVEC_OFF */
    ORC_KO5 = 16       /* FD only: 5th-order Kreiss-Oliger (no KO_ORD_3) */
};

typedef struct {
    int scheme, bc, flags, nz, nvz, gz;
    double z0, z1, v0, v1, CFL;
    double pmo, omega, theta, mu, ko;
} orc_params;

/* Derived per-run constants (src/nuosc.hpp:78-96, 132-143). */
typedef struct {
    int nz, nvz, gz, R; /* R = row length with ghosts */
    double dz, dv, dt, ct, st, pmo, mu, ko;
    double *vz, *Z, *Hm;
} grid_t;

static int grid_init(const orc_params *p, grid_t *g)
{
    g->nz = p->nz;
    g->nvz = p->nvz;
    g->gz = p->gz;
    g->R = p->nz + 2 * p->gz;
    g->dz = (p->z1 - p->z0) / p->nz;
    g->dv = (p->v1 - p->v0) / (p->nvz);
    g->dt = p->CFL * g->dz;
    double th = p->theta * M_PI / 180.0;
    g->ct = (1.0 / 2.0) * p->omega * cos(2.0 * th);
    g->st = (1.0 / 2.0) * p->omega * sin(2.0 * th);
    g->pmo = p->pmo;
    g->mu = p->mu;
    g->ko = p->ko;
    g->vz = (double *)malloc(sizeof(double) * p->nvz);
    g->Z = (double *)malloc(sizeof(double) * p->nz);
    g->Hm = (double *)malloc(sizeof(double) * p->nz);
    if (!g->vz || !g->Z || !g->Hm)
        return -1;
    for (int i = 0; i < p->nvz; i++)
        g->vz[i] = p->v0 + (i + 0.5) * g->dv;
    for (int j = 0; j < p->nz; j++) {
        g->Z[j] = p->z0 + (j + 0.5) * g->dz;
        g->Hm[j] = -1.0e-3 * (g->Z[j] + 600.0);
    }
    return 0;
}

static void grid_free(grid_t *g)
{
    free(g->vz);
    free(g->Z);
    free(g->Hm);
}

static int sgn(double x) { return (0.0 < x) - (x < 0.0); }

/* ------------------------------------------------------------------ */
/* Ghosted field storage: 8 planes of nvz rows of length R.           */

typedef struct {
    double *f[8];
} fields_t;

static int fields_alloc(fields_t *a, size_t n)
{
    for (int k = 0; k < 8; k++) {
        a->f[k] = (double *)calloc(n, sizeof(double));
        if (!a->f[k])
            return -1;
    }
    return 0;
}

static void fields_free(fields_t *a)
{
    for (int k = 0; k < 8; k++)
        free(a->f[k]);
}

#define AT(g, i, j) ((size_t)(i) * (g)->R + (size_t)((j) + (g)->gz))

/* Ghost fill, src/nuosc.hpp:212-278.  The open rule extrapolates the upper
 * ghosts one after another (j ascending), the order of the CPU build. */
static void fill_ghosts(const grid_t *g, int bc, double *u)
{
    const int nz = g->nz, gz = g->gz;
#pragma omp parallel for
    for (int i = 0; i < g->nvz; i++) {
        double *row = u + AT(g, i, 0);
        for (int j = 0; j < gz; j++) {
            if (bc == ORC_PERIODIC) {
                row[-j - 1] = row[nz - j - 1];
                row[nz + j] = row[j];
            } else {
                row[-j - 1] = row[0];
                row[nz + j] = row[nz + j - 1] * 2 - row[nz + j - 2];
            }
        }
    }
}

static void fill_ghosts8(const grid_t *g, int bc, fields_t *a)
{
    for (int k = 0; k < 8; k++)
        fill_ghosts(g, bc, a->f[k]);
}

/* ------------------------------------------------------------------ */
/* WENO7, src/rhs_fv.hpp:18-92.  u points at cell j; the stencil is   */
/* u[-3..3].  smooth() gives the four unscaled smoothness indicators. */

static inline void weno_smooth(const double *u, double si[4])
{
    si[0] = u[0] * (2107 * u[0] - 9402 * u[1] + 7042 * u[2] - 1854 * u[3])
          + u[1] * (11003 * u[1] - 17246 * u[2] + 4642 * u[3])
          + u[2] * (7043 * u[2] - 3882 * u[3]) + 547 * u[3] * u[3];
    si[1] = u[-1] * (547 * u[-1] - 2522 * u[0] + 1922 * u[1] - 494 * u[2])
          + u[0] * (3443 * u[0] - 5966 * u[1] + 1602 * u[2])
          + u[1] * (2843 * u[1] - 1642 * u[2]) + 267 * u[2] * u[2];
    si[2] = u[-2] * (267 * u[-2] - 1642 * u[-1] + 1602 * u[0] - 494 * u[1])
          + u[-1] * (2843 * u[-1] - 5966 * u[0] + 1922 * u[1])
          + u[0] * (3443 * u[0] - 2522 * u[1]) + 547 * u[1] * u[1];
    si[3] = u[-3] * (547 * u[-3] - 3882 * u[-2] + 4642 * u[-1] - 1854 * u[0])
          + u[-2] * (7043 * u[-2] - 17246 * u[-1] + 7042 * u[0])
          + u[-1] * (11003 * u[-1] - 9402 * u[0]) + 2107 * u[0] * u[0];
}

/* Normalised nonlinear weights for linear weights gam[], in the reference's
 * order: w~_k = gam_k / (eps+si_k)^2, w_k = w~_k / (w~_0+w~_1+w~_2+w~_3). */
static inline void weno_weights(const double si[4], const double gam[4], double w[4])
{
    const double eps = 1E-6;
    double a[4];
    for (int k = 0; k < 4; k++) {
        double d = eps + si[k];
        a[k] = gam[k] / (d * d); /* pow(x,2) is x*x in the reference build */
    }
    double s = a[0] + a[1] + a[2] + a[3];
    for (int k = 0; k < 4; k++)
        w[k] = a[k] / s;
}

/* Left-biased value at j+1/2 (the reference's rflux[j]). */
static inline double weno_right_going(const double *u)
{
    static const double gam[4] = {4. / 35., 18. / 35., 12. / 35., 1. / 35.};
    double si[4], w[4];
    weno_smooth(u, si);
    weno_weights(si, gam, w);
    double c0 = (1. / 4.) * u[0] + (13. / 12.) * u[1] - (5. / 12.) * u[2] + (1. / 12.) * u[3];
    double c1 = (-1. / 12.) * u[-1] + (7. / 12.) * u[0] + (7. / 12.) * u[1] - (1. / 12.) * u[2];
    double c2 = (1. / 12.) * u[-2] - (5. / 12.) * u[-1] + (13. / 12.) * u[0] + (1. / 4.) * u[1];
    double c3 = (-1. / 4.) * u[-3] + (13. / 12.) * u[-2] - (23. / 12.) * u[-1] + (25. / 12.) * u[0];
    return w[0] * c0 + w[1] * c1 + w[2] * c2 + w[3] * c3;
}

/* Right-biased value at j-1/2 (the reference's lflux[j]). */
static inline double weno_left_going(const double *u)
{
    static const double gam[4] = {1. / 35., 12. / 35., 18. / 35., 4. / 35.};
    double si[4], w[4];
    weno_smooth(u, si);
    weno_weights(si, gam, w);
    double c0 = (-1. / 4.) * u[3] + (13. / 12.) * u[2] - (23. / 12.) * u[1] + (25. / 12.) * u[0];
    double c1 = (1. / 12.) * u[2] - (5. / 12.) * u[1] + (13. / 12.) * u[0] + (1. / 4.) * u[-1];
    double c2 = (-1. / 12.) * u[1] + (7. / 12.) * u[0] + (7. / 12.) * u[-1] - (1. / 12.) * u[-2];
    double c3 = (1. / 4.) * u[0] + (13. / 12.) * u[-1] - (5. / 12.) * u[-2] + (1. / 12.) * u[-3];
    return w[0] * c0 + w[1] * c1 + w[2] * c2 + w[3] * c3;
}

/* Exposed for the reconstruction-level tests: values at the interfaces
 * j+1/2 (right-going) and j-1/2 (left-going) of a 7-point stencil. */
void orc_weno7(const double *u7, double *right_going, double *left_going)
{
    *right_going = weno_right_going(u7 + 3);
    *left_going = weno_left_going(u7 + 3);
}

/* ------------------------------------------------------------------ */
/* Collective integral.  Brute-force form, src/rhs_fv.hpp:188-232 and */
/* src/rhs_fd.hpp:85-125 (the two are the same arithmetic).           */

typedef struct {
    double ee, xx, exr, exi, bee, bxx, bexr, bexi;
} cell_t;

static inline void collective_brute(const grid_t *g, const fields_t *in, int i, int j,
                                    const cell_t *c, double I[8])
{
    double Iee = 0.0, Ixx = 0.0, Iexr = 0.0, Iexi = 0.0;
    double Ibee = 0.0, Ibxx = 0.0, Ibexr = 0.0, Ibexi = 0.0;
    const double mu = g->mu, vi = g->vz[i];
    for (int k = 0; k < g->nvz; k++) {
        size_t q = AT(g, k, j);
        double eep = in->f[0][q], xxp = in->f[1][q], expr = in->f[2][q], expi = in->f[3][q];
        double beep = in->f[4][q], bxxp = in->f[5][q], bexpr = in->f[6][q], bexpi = in->f[7][q];
        double w = 1.0 - vi * g->vz[k];
        double S = expi + bexpi;  /* Im(rho' + rhobar'), sign-flipped for the barred field */
        double Rr = expr - bexpr; /* Re(rho' - rhobar') */
        double D = eep - xxp - beep + bxxp;
        Iee += 2.0 * mu * w * (c->exr * S - c->exi * Rr);
        Ixx += -2.0 * mu * w * (c->exr * S - c->exi * Rr);
        Iexr += 1.0 * mu * w * ((c->xx - c->ee) * S + c->exi * D);
        Iexi += 1.0 * mu * w * (-(c->xx - c->ee) * Rr - c->exr * D);
        Ibee += 2.0 * mu * w * (-c->bexr * (-expi - bexpi) + c->bexi * Rr);
        Ibxx += -2.0 * mu * w * (-c->bexr * (-expi - bexpi) + c->bexi * Rr);
        Ibexr += 1.0 * mu * w * (-(c->bee - c->bxx) * S - c->bexi * D);
        Ibexi += 1.0 * mu * w * ((c->bee - c->bxx) * (bexpr - expr) + c->bexr * D);
    }
    I[0] = Iee; I[1] = Ixx; I[2] = Iexr; I[3] = Iexi;
    I[4] = Ibee; I[5] = Ibxx; I[6] = Ibexr; I[7] = Ibexi;
}

/* Moment form: sum_k (1 - v_i v_k) f_k = F0 - v_i F1 for f in {S, R, D}.
 * mom[j*6 + {0..5}] = F0S, F1S, F0R, F1R, F0D, F1D.  Each cell's sums run
 * over k = 0..nvz-1 in order; blocks of 64 z cells walk the rows (k outer)
 * so that the z-contiguous rows are read sequentially. */
static void moments(const grid_t *g, const fields_t *in, double *mom)
{
    const int nb = (g->nz + 63) / 64;
#pragma omp parallel for schedule(static)
    for (int blk = 0; blk < nb; blk++) {
        const int j0 = blk * 64, j1 = j0 + 64 < g->nz ? j0 + 64 : g->nz;
        double m[64][6];
        memset(m, 0, sizeof(m));
        for (int k = 0; k < g->nvz; k++) {
            const double v = g->vz[k];
            for (int j = j0; j < j1; j++) {
                size_t q = AT(g, k, j);
                double S = in->f[3][q] + in->f[7][q];
                double Rr = in->f[2][q] - in->f[6][q];
                double D = in->f[0][q] - in->f[1][q] - in->f[4][q] + in->f[5][q];
                double *mm = m[j - j0];
                mm[0] += S; mm[1] += v * S;
                mm[2] += Rr; mm[3] += v * Rr;
                mm[4] += D; mm[5] += v * D;
            }
        }
        for (int j = j0; j < j1; j++)
            for (int q = 0; q < 6; q++)
                mom[(size_t)j * 6 + q] = m[j - j0][q];
    }
}

static inline void collective_moment(const grid_t *g, const double *m, int i, const cell_t *c,
                                     double I[8])
{
    const double mu = g->mu, vi = g->vz[i];
    double S = mu * (m[0] - vi * m[1]);
    double Rr = mu * (m[2] - vi * m[3]);
    double D = mu * (m[4] - vi * m[5]);
    I[0] = 2.0 * (c->exr * S - c->exi * Rr);
    I[1] = -I[0];
    I[2] = (c->xx - c->ee) * S + c->exi * D;
    I[3] = -(c->xx - c->ee) * Rr - c->exr * D;
    I[4] = 2.0 * (c->bexr * S + c->bexi * Rr);
    I[5] = -I[4];
    I[6] = -(c->bee - c->bxx) * S - c->bexi * D;
    I[7] = -(c->bee - c->bxx) * Rr + c->bexr * D;
}

/* ------------------------------------------------------------------ */
/* Right-hand sides.                                                    */

typedef struct {
    const orc_params *p;
    grid_t g;
    fields_t rflux, lflux; /* FV only */
    double *mom;           /* moment form only */
} rhs_ctx;

static void load_cell(const fields_t *in, size_t q, cell_t *c)
{
    c->ee = in->f[0][q]; c->xx = in->f[1][q]; c->exr = in->f[2][q]; c->exi = in->f[3][q];
    c->bee = in->f[4][q]; c->bxx = in->f[5][q]; c->bexr = in->f[6][q]; c->bexi = in->f[7][q];
}

/* FV right-hand side, src/rhs_fv.hpp:98-236.  `in` has filled ghosts. */
static void rhs_fv(rhs_ctx *r, const fields_t *in, fields_t *out, int moment_form)
{
    const grid_t *g = &r->g;
    const int nz = g->nz, nvz = g->nvz, flags = r->p->flags;
    const int advect = !(flags & ORC_ADVEC_OFF);

    if (advect) {
        /* calFlux: both reconstructions for j = -1..nz, then the flux ghost fill. */
        for (int f = 0; f < 8; f++) {
#pragma omp parallel for
            for (int i = 0; i < nvz; i++)
                for (int j = -1; j < nz + 1; j++) {
                    const double *u = &in->f[f][AT(g, i, j)];
                    r->rflux.f[f][AT(g, i, j)] = weno_right_going(u);
                    r->lflux.f[f][AT(g, i, j)] = weno_left_going(u);
                }
        }
        fill_ghosts8(g, r->p->bc, &r->rflux);
        fill_ghosts8(g, r->p->bc, &r->lflux);
    }
    if ((flags & ORC_COLL) && moment_form)
        moments(g, in, r->mom);

#pragma omp parallel for
    for (int i = 0; i < nvz; i++) {
        const double fac = -g->vz[i] / g->dz;
        const int s = sgn(g->vz[i]);
        /* fac*fabs(1+s)/2 and fac*fabs(1-s)/2 select the upwind side exactly. */
        const double sel_r = fac * fabs((double)(1 + s)) / 2;
        const double sel_l = fac * fabs((double)(1 - s)) / 2;
        for (int j = 0; j < nz; j++) {
            size_t q = AT(g, i, j);
            cell_t c;
            load_cell(in, q, &c);
            double o[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (advect)
                for (int f = 0; f < 8; f++)
                    o[f] += sel_r * (r->rflux.f[f][q] - r->rflux.f[f][q - 1])
                          + sel_l * (r->lflux.f[f][q + 1] - r->lflux.f[f][q]);
            if (flags & ORC_VAC) {
                const double st = g->st, ct = g->ct, pmo = g->pmo;
                o[0] += -2.0 * st * c.exi * pmo;
                o[1] += 2.0 * st * c.exi * pmo;
                o[2] += -2.0 * ct * c.exi * pmo;
                o[3] += (2.0 * ct * c.exr + st * (c.ee - c.xx)) * pmo;
                o[4] += -2.0 * st * c.bexi * pmo;
                o[5] += 2.0 * st * c.bexi * pmo;
                o[6] += -2.0 * ct * c.bexi * pmo;
                o[7] += (2.0 * ct * c.bexr + st * (c.bee - c.bxx)) * pmo;
            }
            if (flags & ORC_MAT) {
                const double h = g->Hm[j];
                o[0] += 0.0;
                o[1] += 0.0;
                o[2] += h * c.exi;
                o[3] += -h * c.exr;
                o[4] += 0.0;
                o[5] += 0.0;
                o[6] += -h * c.bexi;
                o[7] += h * c.bexr;
            }
            if (flags & ORC_COLL) {
                double I[8];
                if (moment_form)
                    collective_moment(g, r->mom + (size_t)j * 6, i, &c, I);
                else
                    collective_brute(g, in, i, j, &c, I);
                for (int f = 0; f < 8; f++)
                    o[f] += g->dv * (I[f]);
            }
            for (int f = 0; f < 8; f++)
                out->f[f][q] = o[f];
        }
    }
}

/* FD right-hand side, src/rhs_fd.hpp:6-148 (ADVEC_CENTER_FD; KO order 3
 * unless ORC_KO5).  `in` has filled ghosts. */
static void rhs_fd(rhs_ctx *r, const fields_t *in, fields_t *out, int moment_form)
{
    const grid_t *g = &r->g;
    const int nz = g->nz, nvz = g->nvz, flags = r->p->flags;
    const int ko5 = (flags & ORC_KO5) != 0;

    if ((flags & ORC_COLL) && moment_form)
        moments(g, in, r->mom);

#pragma omp parallel for
    for (int i = 0; i < nvz; i++) {
        const double factor = -g->vz[i] / (12 * g->dz);
        const double ko_eps = ko5 ? -g->ko / g->dz / 64.0 : -g->ko / g->dz / 16.0;
        for (int j = 0; j < nz; j++) {
            size_t q = AT(g, i, j);
            cell_t c;
            load_cell(in, q, &c);
            double o[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            if (flags & ORC_VAC) {
                const double st = g->st, ct = g->ct, pmo = g->pmo;
                o[0] = -pmo * 2 * st * c.exi;
                o[1] = pmo * 2 * st * c.exi;
                o[2] = -pmo * 2 * ct * c.exi;
                o[3] = pmo * (2 * ct * c.exr + st * (c.ee - c.xx));
                o[4] = -pmo * 2 * st * c.bexi;
                o[5] = pmo * 2 * st * c.bexi;
                o[6] = -pmo * 2 * ct * c.bexi;
                o[7] = pmo * (2 * ct * c.bexr + st * (c.bee - c.bxx));
            }
            if (flags & ORC_MAT) {
                const double h = g->Hm[j];
                o[0] += 0.0;
                o[1] += 0.0;
                o[2] += h * c.exi;
                o[3] += -h * c.exr;
                o[4] += 0.0;
                o[5] += 0.0;
                o[6] += -h * c.bexi;
                o[7] += h * c.bexr;
            }
            if (!(flags & ORC_ADVEC_OFF))
                for (int f = 0; f < 8; f++) {
                    const double *x = &in->f[f][q];
                    o[f] += factor * ((x[-2] - x[2]) - 8.0 * (x[-1] - x[1]));
                }
            if (flags & ORC_COLL) {
                double I[8];
                if (moment_form)
                    collective_moment(g, r->mom + (size_t)j * 6, i, &c, I);
                else
                    collective_brute(g, in, i, j, &c, I);
                for (int f = 0; f < 8; f++)
                    o[f] += g->dv * I[f];
            }
            for (int f = 0; f < 8; f++) {
                const double *x = &in->f[f][q];
                double d = ko5 ? (x[-3] + x[3] - 6 * (x[-2] + x[2]) + 15 * (x[-1] + x[1]) - 20 * x[0])
                               : (x[-2] + x[2] - 4 * (x[-1] + x[1]) + 6 * x[0]);
                o[f] += ko_eps * d;
            }
            for (int f = 0; f < 8; f++)
                out->f[f][q] = o[f];
        }
    }
}

static int rhs_ctx_init(rhs_ctx *r, const orc_params *p)
{
    memset(r, 0, sizeof(*r));
    r->p = p;
    int need_gz = p->scheme == ORC_FV ? 4 : ((p->flags & ORC_KO5) ? 3 : 2);
    if (p->nz < 1 || p->nvz < 1 || p->gz < need_gz || p->nz < p->gz)
        return -2;
    if (grid_init(p, &r->g))
        return -1;
    size_t n = (size_t)p->nvz * r->g.R;
    if (p->scheme == ORC_FV && (fields_alloc(&r->rflux, n) || fields_alloc(&r->lflux, n)))
        return -1;
    r->mom = (double *)calloc((size_t)p->nz * 6, sizeof(double));
    return r->mom ? 0 : -1;
}

static void rhs_ctx_free(rhs_ctx *r)
{
    if (r->p->scheme == ORC_FV) {
        fields_free(&r->rflux);
        fields_free(&r->lflux);
    }
    free(r->mom);
    grid_free(&r->g);
}

static void rhs_eval(rhs_ctx *r, fields_t *in, fields_t *out, int moment_form)
{
    fill_ghosts8(&r->g, r->p->bc, in);
    if (r->p->scheme == ORC_FV)
        rhs_fv(r, in, out, moment_form);
    else
        rhs_fd(r, in, out, moment_form);
}

static void to_ghosted(const grid_t *g, const double *state, fields_t *a)
{
    for (int f = 0; f < 8; f++)
        for (int i = 0; i < g->nvz; i++)
            memcpy(&a->f[f][AT(g, i, 0)], state + ((size_t)f * g->nvz + i) * g->nz,
                   sizeof(double) * g->nz);
}

static void from_ghosted(const grid_t *g, const fields_t *a, double *state)
{
    for (int f = 0; f < 8; f++)
        for (int i = 0; i < g->nvz; i++)
            memcpy(state + ((size_t)f * g->nvz + i) * g->nz, &a->f[f][AT(g, i, 0)],
                   sizeof(double) * g->nz);
}

static void set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* One right-hand-side evaluation: out = calRHS(in). */
int orc_rhs(const orc_params *p, const double *in, double *out, int moment_form, int nthreads)
{
    rhs_ctx r;
    int rc = rhs_ctx_init(&r, p);
    if (rc)
        return rc;
    set_threads(nthreads);
    size_t n = (size_t)p->nvz * r.g.R;
    fields_t a, b;
    if (fields_alloc(&a, n) || fields_alloc(&b, n))
        return -1;
    to_ghosted(&r.g, in, &a);
    rhs_eval(&r, &a, &b, moment_form);
    from_ghosted(&r.g, &b, out);
    fields_free(&a);
    fields_free(&b);
    rhs_ctx_free(&r);
    return 0;
}

/* v0 = v1 + a*v2 over interior cells (src/nuosc.hpp:281-298). */
static void axpy(const grid_t *g, fields_t *v0, const fields_t *v1, double a, const fields_t *v2)
{
    for (int f = 0; f < 8; f++) {
#pragma omp parallel for
        for (int i = 0; i < g->nvz; i++)
            for (int j = 0; j < g->nz; j++) {
                size_t q = AT(g, i, j);
                v0->f[f][q] = v1->f[f][q] + a * v2->f[f][q];
            }
    }
}

/* v0 = v1 + a*(v2 + v3) (src/nuosc.hpp:301-318). */
static void axpy2(const grid_t *g, fields_t *v0, const fields_t *v1, double a, const fields_t *v2,
                  const fields_t *v3)
{
    for (int f = 0; f < 8; f++) {
#pragma omp parallel for
        for (int i = 0; i < g->nvz; i++)
            for (int j = 0; j < g->nz; j++) {
                size_t q = AT(g, i, j);
                v0->f[f][q] = v1->f[f][q] + a * (v2->f[f][q] + v3->f[f][q]);
            }
    }
}

static void swapf(fields_t **a, fields_t **b)
{
    fields_t *t = *a;
    *a = *b;
    *b = t;
}

/* nsteps classic RK4 steps in place (src/nuosc.hpp:320-348): the buffer
 * roles y (v_stat), acc (v_rhs), pre (v_pre), cor (v_cor) and their swaps. */
int orc_step(const orc_params *p, double *state, int nsteps, int moment_form, int nthreads)
{
    rhs_ctx r;
    int rc = rhs_ctx_init(&r, p);
    if (rc)
        return rc;
    set_threads(nthreads);
    size_t n = (size_t)p->nvz * r.g.R;
    fields_t bufs[4];
    for (int k = 0; k < 4; k++)
        if (fields_alloc(&bufs[k], n))
            return -1;
    fields_t *y = &bufs[0], *acc = &bufs[1], *pre = &bufs[2], *cor = &bufs[3];
    const grid_t *g = &r.g;
    const double dt = g->dt;
    to_ghosted(g, state, y);
    for (int t = 0; t < nsteps; t++) {
        rhs_eval(&r, y, acc, moment_form);
        axpy(g, pre, y, 0.5 * dt, acc);

        rhs_eval(&r, pre, cor, moment_form);
        axpy(g, acc, acc, 2.0, cor);
        axpy(g, cor, y, 0.5 * dt, cor);
        swapf(&pre, &cor);

        rhs_eval(&r, pre, cor, moment_form);
        axpy(g, acc, acc, 2.0, cor);
        axpy(g, cor, y, dt, cor);
        swapf(&pre, &cor);

        rhs_eval(&r, pre, cor, moment_form);
        axpy2(g, pre, y, 1.0 / 6.0 * dt, cor, acc);
        swapf(&pre, &y);
    }
    from_ghosted(g, y, state);
    for (int k = 0; k < 4; k++)
        fields_free(&bufs[k]);
    rhs_ctx_free(&r);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Initial condition of the reference build, src/initialize.hpp:6-47   */
/* with g/eps/eps_ of src/miscellaneous_funcs.hpp:57-77.                */

double orc_g(double v, double v0, double sigma)
{
    double ex = (v - v0) * (v - v0) / (2.0 * sigma * sigma);
    double N = sigma * sqrt(M_PI / 2.0)
             * (erf((1.0 + v0) / sigma / sqrt(2.0)) + erf((1.0 - v0) / sigma / sqrt(2.0)));
    return exp(-ex) / N;
}

static double pert(double z, double z0, double amp) { return amp * exp(-(z - z0) * (z - z0) / 50.0); }

int orc_init_reference(const orc_params *p, double perturbation_size, double signu, double sigbnu,
                       double alpha, double *G, double *bG, double *state)
{
    grid_t g;
    if (grid_init(p, &g))
        return -1;
    const int nz = p->nz, nvz = p->nvz;
    for (int i = 0; i < nvz; i++)
        for (int j = 0; j < nz; j++) {
            size_t q = (size_t)i * nz + j;
            double Gi = orc_g(g.vz[i], 1.0, signu);
            double bGi = alpha * orc_g(g.vz[i], 1.0, sigbnu);
            double e = pert(g.Z[j], 0.0, perturbation_size);
            double e_ = sqrt(1.0 - e * e);
            G[q] = Gi;
            bG[q] = bGi;
            size_t P = (size_t)nvz * nz;
            state[0 * P + q] = 0.5 * Gi * (1.0 + e_);
            state[1 * P + q] = 0.5 * Gi * (1.0 - e_);
            state[2 * P + q] = 0.5 * Gi * (0.0 + e);
            state[3 * P + q] = -0.0;
            state[4 * P + q] = 0.5 * bGi * (1.0 + e_);
            state[5 * P + q] = 0.5 * bGi * (1.0 - e_);
            state[6 * P + q] = 0.5 * bGi * (0.0 + e);
            state[7 * P + q] = 0.0;
        }
    grid_free(&g);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Diagnostics, src/analysis.hpp:27-185, sequential sums in the        */
/* reference's loop order (v outer, z inner).                          */
/*
 * out[0]  max(dP, dbP)      max |(|P|-|P0|)| over nu and nubar  (dcon)
 * out[1]  <dP>_nu           G-weighted mean                      (dcon)
 * out[2]  <dbP>_nubar       bG-weighted mean                     (dcon)
 * out[3]  |M0|              (cal_Mn, n=0)
 * out[4..6] M0 e1, e2, e3
 * out[7]  <P_ee>            sum ee / sum ee0                     (surv_prob)
 * out[8]  <P_bee>
 * out[9]  trace nu     sum (ee+xx) dz dv
 * out[10] trace nubar  sum (bee+bxx) dz dv
 * out[11] lepton number sum [(ee-xx)-(bee-bxx)] dz dv
 * out[12] 1 if any state value is NaN or Inf, else 0
 */
int orc_diagnostics(const orc_params *p, const double *G, const double *bG, const double *state0,
                    const double *state, double *out)
{
    grid_t g;
    if (grid_init(p, &g))
        return -1;
    const int nz = p->nz, nvz = p->nvz;
    const size_t P = (size_t)nvz * nz;
    const double dz = g.dz, dv = g.dv;
    double dP = 0.0, dbP = 0.0, avdP = 0.0, avbdP = 0.0, Nee0 = 0.0, Nbee0 = 0.0;
    double e1 = 0.0, e2 = 0.0, e3 = 0.0;
    double num = 0, bnum = 0, den = 0, bden = 0, tr = 0, btr = 0, lep = 0;
    int bad = 0;
    for (int i = 0; i < nvz; i++)
        for (int j = 0; j < nz; j++) {
            size_t q = (size_t)i * nz + j;
            const double *s = state, *s0 = state0;
            double P1 = 2.0 * s[2 * P + q] / G[q];
            double P2 = -2.0 * s[3 * P + q] / G[q];
            double P3 = (s[0 * P + q] - s[1 * P + q]) / G[q];
            double nP = fabs(sqrt(P1 * P1 + P2 * P2 + P3 * P3));
            double bP1 = 2.0 * s[6 * P + q] / bG[q];
            double bP2 = 2.0 * s[7 * P + q] / bG[q];
            double bP3 = (s[4 * P + q] - s[5 * P + q]) / bG[q];
            double nbP = fabs(sqrt(bP1 * bP1 + bP2 * bP2 + bP3 * bP3));
            double P01 = 2.0 * s0[2 * P + q] / G[q];
            double P02 = -2.0 * s0[3 * P + q] / G[q];
            double P03 = (s0[0 * P + q] - s0[1 * P + q]) / G[q];
            double nP0 = fabs(sqrt(P01 * P01 + P02 * P02 + P03 * P03));
            double bP01 = 2.0 * s0[6 * P + q] / bG[q];
            double bP02 = 2.0 * s0[7 * P + q] / bG[q];
            double bP03 = (s0[4 * P + q] - s0[5 * P + q]) / bG[q];
            double nbP0 = fabs(sqrt(bP01 * bP01 + bP02 * bP02 + bP03 * bP03));

            e1 += 1.0 * (G[q] * P1 - bG[q] * bP1) * dz * dv;
            e2 += 1.0 * (G[q] * P2 - bG[q] * bP2) * dz * dv;
            e3 += 1.0 * (G[q] * P3 - bG[q] * bP3) * dz * dv;

            double a = fabs(nP - nP0), b = fabs(nbP - nbP0);
            dP = (dP >= a) ? dP : a;
            dbP = (dbP >= b) ? dbP : b;
            avdP += a * G[q];
            avbdP += b * bG[q];
            Nee0 += G[q];
            Nbee0 += bG[q];

            num += s[0 * P + q] * dz * dv;
            bnum += s[4 * P + q] * dz * dv;
            den += s0[0 * P + q] * dz * dv;
            bden += s0[4 * P + q] * dz * dv;

            tr += (s[0 * P + q] + s[1 * P + q]) * dz * dv;
            btr += (s[4 * P + q] + s[5 * P + q]) * dz * dv;
            lep += ((s[0 * P + q] - s[1 * P + q]) - (s[4 * P + q] - s[5 * P + q])) * dz * dv;
            for (int f = 0; f < 8; f++)
                if (!isfinite(s[f * P + q]))
                    bad = 1;
        }
    out[0] = (dP >= dbP) ? dP : dbP;
    out[1] = avdP / Nee0;
    out[2] = avbdP / Nbee0;
    out[3] = sqrt(e1 * e1 + e2 * e2 + e3 * e3);
    out[4] = e1;
    out[5] = e2;
    out[6] = e3;
    out[7] = num / den;
    out[8] = bnum / bden;
    out[9] = tr;
    out[10] = btr;
    out[11] = lep;
    out[12] = bad;
    grid_free(&g);
    return 0;
}

/* Grid constants as the reference derives them: out = {dz, dv, dt, ct, st}. */
int orc_grid(const orc_params *p, double *out, double *vz, double *Z)
{
    grid_t g;
    if (grid_init(p, &g))
        return -1;
    out[0] = g.dz; out[1] = g.dv; out[2] = g.dt; out[3] = g.ct; out[4] = g.st;
    if (vz)
        memcpy(vz, g.vz, sizeof(double) * p->nvz);
    if (Z)
        memcpy(Z, g.Z, sizeof(double) * p->nz);
    grid_free(&g);
    return 0;
}
