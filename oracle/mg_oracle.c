/*
 * CPU ORACLE -- test infrastructure only (see oracle.h).
 *
 * Restates, loop for loop, the 3D geometric multigrid of SARAS:
 *   lib/poisson/poisson.cc     poisson::poisson 62-83, mgSolve 100-205, vCycle 219-280,
 *                              setStagBounds 326-362, setCoefficients 399-439, copyStaggrDerivs 453-496
 *   lib/poisson/poisson_d3.cc  computeResidual 90-111, smooth 114-187, solve 190-247, coarsen 250-298,
 *                              prolong 301-388, computeError 391-459, initDirichlet 513-572,
 *                              imposeBC 575-668, updatePads 671-686 (single-rank self exchange)
 *   lib/grid/grid.cc           createUniformGrid 370-422, createTanHypGrid 442-519
 * plus two additions the GPU path needs a checker for: the red-black Gauss-Seidel order and the
 * range projection of a singular right-hand side.
 *
 * Parity pins (tests/test_oracle_mg.py): the reference's own known-answer tests testProlong
 * (poisson_d3.cc:689-719), testTransfer (722-777) and testPeriodic (780-857); exactness of the
 * discrete Laplacian on r^2/6 (poisson.cc:153-204 with hydro_d3.cc:332-334); closed-form
 * restriction weights and restriction of constants; Jacobi/GS sweep identities by brute force;
 * dense direct solve of the same discrete system on small grids; O(h^2) error against the
 * analytic solution of the sin-manufactured Dirichlet problem.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NMET 6

typedef struct {
    int nx, ny, nz;                       /* core nodes: indices 0..n-1 */
    long sx, sy;                          /* padded strides (k stride is 1) */
    long size;                            /* padded element count */
    double hx, hy, hz, hxhy, hyhz, hzhx, hxhyhz;
    double *met[NMET];                    /* padded 1D metrics, element 0 = index -1 */
} level_t;

struct or_mg {
    or_params p;
    int depth;
    int vLevel;
    int zeroBC;
    int maxCount;
    level_t *lv;
    double **P, **R, **T, **S;            /* pressureData, residualData, tmpDataArray, smoothedPres */
    double *wall[6];                      /* x lo/hi, y lo/hi, z lo/hi (level 0, padded 2D) */
};

static inline long IX(const level_t *L, int i, int j, int k) {
    return (long)(i + 1) * L->sx + (long)(j + 1) * L->sy + (k + 1);
}

/* ------------------------------------------------------------------------------------------ */
/* grid.cc                                                                                       */
/* ------------------------------------------------------------------------------------------ */

/* grid.cc:155-164 sizeArray; 376-378 (uniform), 461-468 (tanh, core indices only) */
void or_grid_staggr(int n, double len, int grid_type, double beta,
                    double *x, double *xi_x, double *xixx, double *xix2) {
    double dxi = 1.0 / (double)(n - 1);
    for (int i = -1; i <= n; i++) {
        double xi = (double)i * dxi;
        x[i + 1] = len * xi;
        xi_x[i + 1] = 1.0 / len;
        xixx[i + 1] = 0.0;
        xix2[i + 1] = pow(1.0 / len, 2.0);
    }
    if (grid_type == 2) {
        double tb = tanh(beta);
        for (int i = 0; i < n; i++) {
            double xi = (double)i * dxi;
            double xs = len * (1.0 - tanh(beta * (1.0 - 2.0 * xi)) / tb) / 2.0;
            x[i + 1] = xs;
            xi_x[i + 1] = tb / (beta * len * (1.0 - pow((1.0 - 2.0 * xs / len) * tb, 2)));
            xixx[i + 1] = -4.0 * pow(tb, 3) * (1.0 - 2.0 * xs / len) /
                          (beta * len * len * pow(1.0 - pow(tb * (1.0 - 2.0 * xs / len), 2), 2));
            xix2[i + 1] = pow(xi_x[i + 1], 2.0);
        }
    }
}

/* grid.cc:381-383 (uniform), 471-477 (tanh); n = number of nodes, faces = n-1 */
void or_grid_colloc(int n, double len, int grid_type, double beta,
                    double *x, double *xi_x, double *xixx, double *xix2) {
    double dxi = 1.0 / (double)(n - 1);
    int nc = n - 1;
    for (int i = -1; i <= nc; i++) {
        double a = (double)i * dxi, b = (double)(i + 1) * dxi;
        x[i + 1] = len * (a + b) / 2.0;
        xi_x[i + 1] = 1.0 / len;
        xixx[i + 1] = 0.0;
        xix2[i + 1] = pow(1.0 / len, 2.0);
    }
    if (grid_type == 2) {
        double tb = tanh(beta);
        for (int i = 0; i < nc; i++) {
            double a = (double)i * dxi, b = (double)(i + 1) * dxi;
            double xs = len * (1.0 - tanh(beta * (1.0 - (a + b))) / tb) / 2.0;
            x[i + 1] = xs;
            xi_x[i + 1] = tb / (beta * len * (1.0 - pow((1.0 - 2.0 * xs / len) * tb, 2)));
            xixx[i + 1] = -4.0 * pow(tb, 3) * (1.0 - 2.0 * xs / len) /
                          (beta * len * len * pow(1.0 - pow(tb * (1.0 - 2.0 * xs / len), 2), 2));
            xix2[i + 1] = pow(xi_x[i + 1], 2.0);
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* construction: poisson.cc:62-83, 326-362, 399-439, 453-496, 293-316                           */
/* ------------------------------------------------------------------------------------------ */

or_mg *or_mg_create(const or_params *p,
                    const double *xix2, const double *xixx,
                    const double *ety2, const double *etyy,
                    const double *ztz2, const double *ztzz,
                    const double *xWallLo, const double *xWallHi,
                    const double *yWallLo, const double *yWallHi,
                    const double *zWallLo, const double *zWallHi) {
    or_mg *m = (or_mg *)calloc(1, sizeof(or_mg));
    m->p = *p;
    int d = p->vcDepth;
    m->depth = d;
    m->lv = (level_t *)calloc((size_t)d + 1, sizeof(level_t));
    m->P = (double **)calloc((size_t)d + 1, sizeof(double *));
    m->R = (double **)calloc((size_t)d + 1, sizeof(double *));
    m->T = (double **)calloc((size_t)d + 1, sizeof(double *));
    m->S = (double **)calloc((size_t)d + 1, sizeof(double *));
    const double *fine[NMET] = {xix2, xixx, ety2, etyy, ztz2, ztzz};
    int n0[3] = {(1 << p->xInd) + 1, (1 << p->yInd) + 1, (1 << p->zInd) + 1};
    for (int l = 0; l <= d; l++) {
        level_t *L = &m->lv[l];
        int stride = 1 << l;
        L->nx = (1 << (p->xInd - l)) + 1;   /* mgSizeArray(localSizeIndex - l) */
        L->ny = (1 << (p->yInd - l)) + 1;
        L->nz = (1 << (p->zInd - l)) + 1;
        L->sy = L->nz + 2;
        L->sx = (long)(L->ny + 2) * L->sy;
        L->size = (long)(L->nx + 2) * L->sx;
        double s = (double)stride;
        L->hx = s * p->dXi;
        L->hy = s * p->dEt;
        L->hz = s * p->dZt;
        L->hxhy = pow(s, 4.0) * pow(p->dXi, 2.0) * pow(p->dEt, 2.0);
        L->hyhz = pow(s, 4.0) * pow(p->dEt, 2.0) * pow(p->dZt, 2.0);
        L->hzhx = pow(s, 4.0) * pow(p->dZt, 2.0) * pow(p->dXi, 2.0);
        L->hxhyhz = pow(s, 6.0) * pow(p->dXi, 2.0) * pow(p->dEt, 2.0) * pow(p->dZt, 2.0);
        int nl[3] = {L->nx, L->ny, L->nz};
        for (int q = 0; q < NMET; q++) {
            int dim = q / 2;
            int n = nl[dim];
            L->met[q] = (double *)calloc((size_t)n + 2, sizeof(double));
            /* core 0..n-1 sampled from the finest core at stride 2^l; pads stay 0 (poisson.cc:467-494) */
            for (int i = 0; i < n; i++) {
                int src = i * stride;
                (void)n0;
                L->met[q][i + 1] = fine[q][src + 1];
            }
        }
        m->P[l] = (double *)calloc((size_t)L->size, sizeof(double));
        m->R[l] = (double *)calloc((size_t)L->size, sizeof(double));
        m->T[l] = (double *)calloc((size_t)L->size, sizeof(double));
        m->S[l] = (double *)calloc((size_t)L->size, sizeof(double));
    }
    /* poisson.cc:359-361: maxCount = product of (N_c + 1) over the coarsest level */
    level_t *C = &m->lv[d];
    m->maxCount = (C->nx + 1) * (C->ny + 1) * (C->nz + 1);

    const double *walls[6] = {xWallLo, xWallHi, yWallLo, yWallHi, zWallLo, zWallHi};
    level_t *F = &m->lv[0];
    long wsz[3] = {(long)(F->ny + 2) * (F->nz + 2), (long)(F->nx + 2) * (F->nz + 2),
                   (long)(F->nx + 2) * (F->ny + 2)};
    for (int w = 0; w < 6; w++) {
        if (walls[w]) {
            m->wall[w] = (double *)malloc(sizeof(double) * (size_t)wsz[w / 2]);
            memcpy(m->wall[w], walls[w], sizeof(double) * (size_t)wsz[w / 2]);
        }
    }
    return m;
}

void or_mg_destroy(or_mg *m) {
    if (!m) return;
    for (int l = 0; l <= m->depth; l++) {
        for (int q = 0; q < NMET; q++) free(m->lv[l].met[q]);
        free(m->P[l]); free(m->R[l]); free(m->T[l]); free(m->S[l]);
    }
    for (int w = 0; w < 6; w++) free(m->wall[w]);
    free(m->P); free(m->R); free(m->T); free(m->S); free(m->lv);
    free(m);
}

double *or_mg_array(or_mg *m, int which, int level) {
    switch (which) {
        case 0: return m->P[level];
        case 1: return m->R[level];
        case 2: return m->T[level];
        default: return m->S[level];
    }
}

void or_mg_dims(const or_mg *m, int level, int *nx, int *ny, int *nz) {
    *nx = m->lv[level].nx; *ny = m->lv[level].ny; *nz = m->lv[level].nz;
}
void or_mg_set_level(or_mg *m, int level) { m->vLevel = level; }
int  or_mg_get_level(const or_mg *m) { return m->vLevel; }
void or_mg_set_zero_bc(or_mg *m, int zeroBC) { m->zeroBC = zeroBC; }
const double *or_mg_metric(const or_mg *m, int which, int level) { return m->lv[level].met[which]; }
int or_mg_max_count(const or_mg *m) { return m->maxCount; }
double or_mg_coef(const or_mg *m, int which, int level) {
    const level_t *L = &m->lv[level];
    switch (which) {
        case 0: return L->hx;   case 1: return L->hy;   case 2: return L->hz;
        case 3: return L->hxhy; case 4: return L->hyhz; case 5: return L->hzhx;
        default: return L->hxhyhz;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* pads and boundary conditions                                                                 */
/* ------------------------------------------------------------------------------------------ */

/* poisson_d3.cc:671-686 with faceRanks = self along periodic x/y and MPI_PROC_NULL otherwise
 * (hydro.cc:174-212): the x exchange copies whole padded planes, then the y exchange copies
 * whole padded x-z slabs (which by then carry the updated x pads).                              */
static void update_pads_arr(const or_mg *m, double *a, int l) {
    const level_t *L = &m->lv[l];
    int Mx = L->nx - 1, My = L->ny - 1;
    if (m->p.xPer) {
        for (int j = -1; j <= L->ny; j++)
            for (int k = -1; k <= L->nz; k++) {
                a[IX(L, -1, j, k)] = a[IX(L, Mx - 1, j, k)];
                a[IX(L, Mx + 1, j, k)] = a[IX(L, 1, j, k)];
            }
    }
    if (m->p.yPer) {
        for (int i = -1; i <= L->nx; i++)
            for (int k = -1; k <= L->nz; k++) {
                a[IX(L, i, -1, k)] = a[IX(L, i, My - 1, k)];
                a[IX(L, i, My + 1, k)] = a[IX(L, i, 1, k)];
            }
    }
}

void or_mg_update_pads(or_mg *m, int which) {
    update_pads_arr(m, or_mg_array(m, which, m->vLevel), m->vLevel);
}

/* poisson_d3.cc:575-668 */
void or_mg_impose_bc(or_mg *m) {
    int l = m->vLevel;
    const level_t *L = &m->lv[l];
    double *a = m->P[l];
    int Mx = L->nx - 1, My = L->ny - 1, Mz = L->nz - 1;
    update_pads_arr(m, a, l);
    if (!m->p.xPer) {
        for (int j = -1; j <= L->ny; j++)
            for (int k = -1; k <= L->nz; k++) {
                if (m->p.testPoisson) {
                    if (m->zeroBC) {
                        a[IX(L, -1, j, k)] = -a[IX(L, 1, j, k)];
                        a[IX(L, Mx + 1, j, k)] = -a[IX(L, Mx - 1, j, k)];
                    } else {
                        long w = (long)(j + 1) * (L->nz + 2) + (k + 1);
                        a[IX(L, -1, j, k)] = m->wall[0] ? m->wall[0][w] : 0.0;
                        a[IX(L, Mx + 1, j, k)] = m->wall[1] ? m->wall[1][w] : 0.0;
                    }
                } else {
                    a[IX(L, -1, j, k)] = a[IX(L, 1, j, k)];
                    a[IX(L, Mx + 1, j, k)] = a[IX(L, Mx - 1, j, k)];
                }
            }
    }
    if (!m->p.yPer) {
        for (int i = -1; i <= L->nx; i++)
            for (int k = -1; k <= L->nz; k++) {
                if (m->p.testPoisson) {
                    if (m->zeroBC) {
                        a[IX(L, i, -1, k)] = -a[IX(L, i, 1, k)];
                        a[IX(L, i, My + 1, k)] = -a[IX(L, i, My - 1, k)];
                    } else {
                        long w = (long)(i + 1) * (L->nz + 2) + (k + 1);
                        a[IX(L, i, -1, k)] = m->wall[2] ? m->wall[2][w] : 0.0;
                        a[IX(L, i, My + 1, k)] = m->wall[3] ? m->wall[3][w] : 0.0;
                    }
                } else {
                    a[IX(L, i, -1, k)] = a[IX(L, i, 1, k)];
                    a[IX(L, i, My + 1, k)] = a[IX(L, i, My - 1, k)];
                }
            }
    }
    for (int i = -1; i <= L->nx; i++)
        for (int j = -1; j <= L->ny; j++) {
            if (m->p.zPer) {
                a[IX(L, i, j, -1)] = a[IX(L, i, j, Mz - 1)];
                a[IX(L, i, j, Mz + 1)] = a[IX(L, i, j, 1)];
            } else if (m->p.testPoisson) {
                if (m->zeroBC) {
                    a[IX(L, i, j, -1)] = -a[IX(L, i, j, 1)];
                    a[IX(L, i, j, Mz + 1)] = -a[IX(L, i, j, Mz - 1)];
                } else {
                    long w = (long)(i + 1) * (L->ny + 2) + (j + 1);
                    a[IX(L, i, j, -1)] = m->wall[4] ? m->wall[4][w] : 0.0;
                    a[IX(L, i, j, Mz + 1)] = m->wall[5] ? m->wall[5][w] : 0.0;
                }
            } else {
                a[IX(L, i, j, -1)] = a[IX(L, i, j, 1)];
                a[IX(L, i, j, Mz + 1)] = a[IX(L, i, j, Mz - 1)];
            }
        }
}

/* ------------------------------------------------------------------------------------------ */
/* stencils                                                                                     */
/* ------------------------------------------------------------------------------------------ */

/* Laplacian as written in poisson_d3.cc:99-105 (and 209-216, 409-414) */
static inline double lap(const level_t *L, const double *u, int i, int j, int k, double hx, double hy, double hz) {
    long c = IX(L, i, j, k);
    const double *xix2 = L->met[0], *xixx = L->met[1], *ety2 = L->met[2], *etyy = L->met[3],
                 *ztz2 = L->met[4], *ztzz = L->met[5];
    return xix2[i + 1] * (u[c + L->sx] - 2.0 * u[c] + u[c - L->sx]) / (hx * hx) +
           xixx[i + 1] * (u[c + L->sx] - u[c - L->sx]) / (2.0 * hx) +
           ety2[j + 1] * (u[c + L->sy] - 2.0 * u[c] + u[c - L->sy]) / (hy * hy) +
           etyy[j + 1] * (u[c + L->sy] - u[c - L->sy]) / (2.0 * hy) +
           ztz2[k + 1] * (u[c + 1] - 2.0 * u[c] + u[c - 1]) / (hz * hz) +
           ztzz[k + 1] * (u[c + 1] - u[c - 1]) / (2.0 * hz);
}

/* Gauss-Seidel / Jacobi update as written in poisson_d3.cc:140-147 */
static inline double gs_value(const level_t *L, const double *u, const double *f, int i, int j, int k) {
    long c = IX(L, i, j, k);
    const double *xix2 = L->met[0], *xixx = L->met[1], *ety2 = L->met[2], *etyy = L->met[3],
                 *ztz2 = L->met[4], *ztzz = L->met[5];
    return (L->hyhz * xix2[i + 1] * (u[c + L->sx] + u[c - L->sx]) * 2.0 +
            L->hyhz * xixx[i + 1] * (u[c + L->sx] - u[c - L->sx]) * L->hx +
            L->hzhx * ety2[j + 1] * (u[c + L->sy] + u[c - L->sy]) * 2.0 +
            L->hzhx * etyy[j + 1] * (u[c + L->sy] - u[c - L->sy]) * L->hy +
            L->hxhy * ztz2[k + 1] * (u[c + 1] + u[c - 1]) * 2.0 +
            L->hxhy * ztzz[k + 1] * (u[c + 1] - u[c - 1]) * L->hz -
            2.0 * L->hxhyhz * f[c]) /
           (4.0 * (L->hyhz * xix2[i + 1] + L->hzhx * ety2[j + 1] + L->hxhy * ztz2[k + 1]));
}

/* Addition (padMode 1): residual pads that follow the ghost rule of each direction -- mirror for
 * Neumann, wrap for periodic z, zero for Dirichlet -- so that full weighting equals the Galerkin
 * transpose of trilinear prolongation for the full-equation boundary rows.                      */
static void consistent_pads(const or_mg *m, double *a, int l) {
    const level_t *L = &m->lv[l];
    int Mx = L->nx - 1, My = L->ny - 1, Mz = L->nz - 1;
    int dir = m->p.testPoisson;
    if (!m->p.xPer && !dir)
        for (int j = -1; j <= L->ny; j++)
            for (int k = -1; k <= L->nz; k++) {
                a[IX(L, -1, j, k)] = a[IX(L, 1, j, k)];
                a[IX(L, Mx + 1, j, k)] = a[IX(L, Mx - 1, j, k)];
            }
    if (!m->p.yPer && !dir)
        for (int i = -1; i <= L->nx; i++)
            for (int k = -1; k <= L->nz; k++) {
                a[IX(L, i, -1, k)] = a[IX(L, i, 1, k)];
                a[IX(L, i, My + 1, k)] = a[IX(L, i, My - 1, k)];
            }
    if (m->p.zPer || !dir)
        for (int i = -1; i <= L->nx; i++)
            for (int j = -1; j <= L->ny; j++) {
                a[IX(L, i, j, -1)] = a[IX(L, i, j, m->p.zPer ? Mz - 1 : 1)];
                a[IX(L, i, j, Mz + 1)] = a[IX(L, i, j, m->p.zPer ? 1 : Mz - 1)];
            }
}

/* poisson_d3.cc:90-111 */
void or_mg_compute_residual(or_mg *m) {
    int l = m->vLevel;
    const level_t *L = &m->lv[l];
    double *t = m->T[l];
    const double *u = m->P[l], *f = m->R[l];
    memset(t, 0, sizeof(double) * (size_t)L->size);
    for (int i = 0; i < L->nx; i++)
        for (int j = 0; j < L->ny; j++)
            for (int k = 0; k < L->nz; k++)
                t[IX(L, i, j, k)] = f[IX(L, i, j, k)] - lap(L, u, i, j, k, L->hx, L->hy, L->hz);
    update_pads_arr(m, t, l);
    if (m->p.padMode == 1) consistent_pads(m, t, l);
}

static void sweep_lex(or_mg *m, int l) {
    const level_t *L = &m->lv[l];
    double *u = m->P[l];
    const double *f = m->R[l];
    for (int i = 0; i < L->nx; i++)
        for (int j = 0; j < L->ny; j++)
            for (int k = 0; k < L->nz; k++)
                u[IX(L, i, j, k)] = gs_value(L, u, f, i, j, k);
}

static void sweep_color(or_mg *m, int l, int color) {
    const level_t *L = &m->lv[l];
    double *u = m->P[l];
    const double *f = m->R[l];
    for (int i = 0; i < L->nx; i++)
        for (int j = 0; j < L->ny; j++)
            for (int k = ((i + j + color) & 1); k < L->nz; k += 2)
                u[IX(L, i, j, k)] = gs_value(L, u, f, i, j, k);
}

/* one red-black sweep: ghosts refreshed before each colour */
static void sweep_rb(or_mg *m, int l) {
    or_mg_impose_bc(m);
    sweep_color(m, l, 0);
    or_mg_impose_bc(m);
    sweep_color(m, l, 1);
}

/* poisson_d3.cc:114-187 */
void or_mg_smooth(or_mg *m, int count) {
    int l = m->vLevel;
    const level_t *L = &m->lv[l];
    memset(m->T[l], 0, sizeof(double) * (size_t)L->size);
    for (int n = 0; n < count; n++) {
        if (m->p.smoother == OR_SMOOTH_RB) {
            sweep_rb(m, l);
        } else if (m->p.smoother == OR_SMOOTH_LEX) {
            or_mg_impose_bc(m);
            sweep_lex(m, l);
        } else {
            or_mg_impose_bc(m);
            double *u = m->P[l], *t = m->T[l];
            const double *f = m->R[l];
            for (int i = 0; i < L->nx; i++)
                for (int j = 0; j < L->ny; j++)
                    for (int k = 0; k < L->nz; k++)
                        t[IX(L, i, j, k)] = gs_value(L, u, f, i, j, k);
            /* swap(tmpDataArray, pressureData): the whole hierarchies exchange handles */
            double **tmp = m->T; m->T = m->P; m->P = tmp;
        }
    }
    or_mg_impose_bc(m);
}

static double max_abs_residual(const or_mg *m, int l) {
    const level_t *L = &m->lv[l];
    const double *u = m->P[l], *f = m->R[l];
    double localMax = -1.0e-10;
    for (int i = 0; i < L->nx; i++)
        for (int j = 0; j < L->ny; j++)
            for (int k = 0; k < L->nz; k++) {
                double v = fabs(f[IX(L, i, j, k)] - lap(L, u, i, j, k, L->hx, L->hy, L->hz));
                if (v > localMax) localMax = v;
            }
    return localMax;
}

/* poisson_d3.cc:190-247 (the red-black order replaces the lexicographic sweep in OR_SMOOTH_RB) */
void or_mg_solve_coarsest(or_mg *m) {
    int l = m->vLevel;
    int iterCount = 0;
    while (1) {
        if (m->p.smoother == OR_SMOOTH_RB || m->p.smoother == OR_SMOOTH_JACOBI_RB) {
            sweep_rb(m, l);
        } else {
            or_mg_impose_bc(m);
            sweep_lex(m, l);
        }
        double gmax = max_abs_residual(m, l);
        if (gmax < m->p.mgTolerance) break;
        iterCount += 1;
        if (iterCount > m->maxCount) break;
    }
    or_mg_impose_bc(m);
}

/* poisson_d3.cc:250-298 */
void or_mg_coarsen(or_mg *m) {
    int pl = m->vLevel;
    m->vLevel += 1;
    int l = m->vLevel;
    const level_t *F = &m->lv[pl], *C = &m->lv[l];
    const double *t = m->T[pl];
    double *r = m->R[l];
    for (int i = 0; i < C->nx; i++) {
        int i2 = i * 2;
        for (int j = 0; j < C->ny; j++) {
            int j2 = j * 2;
            for (int k = 0; k < C->nz; k++) {
                int k2 = k * 2;
#define TT(a, b, c) t[IX(F, i2 + (a), j2 + (b), k2 + (c))]
                double facePoints = (TT(1, 0, 0) + TT(-1, 0, 0) + TT(0, 1, 0) + TT(0, -1, 0) +
                                     TT(0, 0, 1) + TT(0, 0, -1)) * 0.0625;
                double edgePoints = (TT(1, 1, 0) + TT(1, -1, 0) + TT(-1, -1, 0) + TT(-1, 1, 0) +
                                     TT(0, 1, 1) + TT(0, -1, 1) + TT(0, -1, -1) + TT(0, 1, -1) +
                                     TT(1, 0, 1) + TT(1, 0, -1) + TT(-1, 0, -1) + TT(-1, 0, 1)) * 0.03125;
                double vertPoints = (TT(1, 1, 1) + TT(1, 1, -1) + TT(1, -1, 1) + TT(-1, 1, 1) +
                                     TT(1, -1, -1) + TT(-1, 1, -1) + TT(-1, -1, 1) + TT(-1, -1, -1)) * 0.015625;
                r[IX(C, i, j, k)] = facePoints + edgePoints + vertPoints + TT(0, 0, 0) * 0.125;
#undef TT
            }
        }
    }
}

/* poisson_d3.cc:301-388 */
void or_mg_prolong(or_mg *m) {
    int pl = m->vLevel;
    m->vLevel -= 1;
    int l = m->vLevel;
    const level_t *F = &m->lv[l], *C = &m->lv[pl];
    double *u = m->P[l];
    const double *e = m->P[pl];
    memset(u, 0, sizeof(double) * (size_t)F->size);
    for (int i = 0; i < F->nx; i++) {
        int i2 = i / 2, io = i & 1;
        for (int j = 0; j < F->ny; j++) {
            int j2 = j / 2, jo = j & 1;
            for (int k = 0; k < F->nz; k++) {
                int k2 = k / 2, ko = k & 1;
#define EE(a, b, c) e[IX(C, i2 + (a), j2 + (b), k2 + (c))]
                double v;
                if (io && jo && ko)
                    v = (EE(0, 0, 0) + EE(1, 0, 0) + EE(0, 1, 0) + EE(0, 0, 1) + EE(1, 1, 0) + EE(1, 0, 1) +
                         EE(0, 1, 1) + EE(1, 1, 1)) / 8.0;
                else if (io && jo)
                    v = (EE(0, 0, 0) + EE(0, 1, 0) + EE(1, 0, 0) + EE(1, 1, 0)) / 4.0;
                else if (io && ko)
                    v = (EE(0, 0, 0) + EE(0, 0, 1) + EE(1, 0, 0) + EE(1, 0, 1)) / 4.0;
                else if (io)
                    v = (EE(0, 0, 0) + EE(1, 0, 0)) / 2.0;
                else if (jo && ko)
                    v = (EE(0, 0, 0) + EE(0, 0, 1) + EE(0, 1, 0) + EE(0, 1, 1)) / 4.0;
                else if (jo)
                    v = (EE(0, 0, 0) + EE(0, 1, 0)) / 2.0;
                else if (ko)
                    v = (EE(0, 0, 0) + EE(0, 0, 1)) / 2.0;
                else
                    v = EE(0, 0, 0);
#undef EE
                u[IX(F, i, j, k)] = v;
            }
        }
    }
}

/* poisson_d3.cc:391-459 (single rank: the Allreduce calls are identities) */
double or_mg_compute_error(or_mg *m, int normOrder) {
    const level_t *L = &m->lv[0];
    const level_t *V = &m->lv[m->vLevel];
    const double *u = m->P[0], *f = m->R[0];
    double num = 0.0, den = 0.0;
    long cnt = 0;
    for (int i = 0; i < L->nx; i++)
        for (int j = 0; j < L->ny; j++)
            for (int k = 0; k < L->nz; k++) {
                double v = fabs(lap(L, u, i, j, k, V->hx, V->hy, V->hz) - f[IX(L, i, j, k)]);
                if (normOrder == 1) {
                    if (v > num) num = v;
                } else if (normOrder == 2) {
                    num += v * v;
                    den += f[IX(L, i, j, k)] * f[IX(L, i, j, k)];
                    cnt += 1;
                }
            }
    if (normOrder == 1) {
        den = 0.0;
        for (long q = 0; q < L->size; q++)
            if (fabs(f[q]) > den) den = fabs(f[q]);
        return den != 0.0 ? num / den : num;
    }
    if (normOrder == 2) {
        double c = (double)cnt;
        return den != 0.0 ? sqrt(num / c) / sqrt(den / c) : sqrt(num / c);
    }
    return 0.0;
}

static int has_dirichlet(const or_params *p) {
    return p->testPoisson && !(p->xPer && p->yPer && p->zPer);
}
static void project_level(const or_mg *m, int l, double *rhs);

/* poisson.cc:219-280 */
void or_mg_vcycle(or_mg *m) {
    const or_params *p = &m->p;
    m->vLevel = 0;
    m->zeroBC = 0;
    or_mg_smooth(m, p->preSmooth);
    m->zeroBC = 1;
    for (int i = 0; i < p->vcDepth; i++) {
        or_mg_compute_residual(m);
        int l = m->vLevel;
        memcpy(m->S[l], m->P[l], sizeof(double) * (size_t)m->lv[l].size);
        or_mg_coarsen(m);
        memset(m->P[m->vLevel], 0, sizeof(double) * (size_t)m->lv[m->vLevel].size);
        if (m->vLevel == p->vcDepth) {
            if (p->coarseProject && !has_dirichlet(p)) project_level(m, m->vLevel, m->R[m->vLevel]);
            or_mg_solve_coarsest(m);
        }
        else or_mg_smooth(m, p->preSmooth);
    }
    for (int i = 0; i < p->vcDepth; i++) {
        or_mg_prolong(m);
        int l = m->vLevel;
        double *u = m->P[l];
        const double *s = m->S[l];
        for (long q = 0; q < m->lv[l].size; q++) u[q] += s[q];
        m->zeroBC = (l == 0) ? 0 : 1;
        or_mg_smooth(m, p->postSmooth);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* range projection (addition: the reference subtracts a padded-array mean instead)             */
/* ------------------------------------------------------------------------------------------ */

/* Left null vector of the 1D three-point operator of one direction at the finest level.
 * Periodic (uniform metric): trapezoid weights, 1/2 at the duplicated end nodes.
 * Mirror (Neumann) ghosts: the operator has zero row sums, so w satisfies the balance
 * w_k c_k = w_{k+1} a_{k+1} with a, c the sub/super-diagonal (mirror folded into row 0 / n-1). */
void or_null_vector_1d(int n, int periodic, double h, const double *d2, const double *d1, double *w) {
    if (periodic) {
        for (int k = 0; k < n; k++) w[k] = 1.0;
        w[0] = 0.5;
        w[n - 1] = 0.5;
        return;
    }
    /* d2, d1: core metric arrays (length n) of the second- and first-derivative terms */
    double h2 = h * h;
    w[0] = 1.0;
    for (int k = 0; k < n - 1; k++) {
        double ck = d2[k] / h2 + d1[k] / (2.0 * h);
        double ak1 = d2[k + 1] / h2 - d1[k + 1] / (2.0 * h);
        if (k == 0) ck += d2[0] / h2 - d1[0] / (2.0 * h);          /* u(-1) = u(1) folded into row 0 */
        if (k + 1 == n - 1) ak1 += d2[n - 1] / h2 + d1[n - 1] / (2.0 * h); /* u(n) = u(n-2) into row n-1 */
        w[k + 1] = w[k] * ck / ak1;
    }
}

static void project_level(const or_mg *m, int l, double *rhs);

void or_mg_project_rhs(const or_mg *m, double *rhs) { project_level(m, 0, rhs); }

static void project_level(const or_mg *m, int l, double *rhs) {
    const level_t *L = &m->lv[l];
    int n[3] = {L->nx, L->ny, L->nz};
    int per[3] = {m->p.xPer, m->p.yPer, m->p.zPer};
    double h[3] = {L->hx, L->hy, L->hz};
    double *w[3];
    for (int d = 0; d < 3; d++) {
        w[d] = (double *)malloc(sizeof(double) * (size_t)n[d]);
        or_null_vector_1d(n[d], per[d], h[d], L->met[2 * d] + 1, L->met[2 * d + 1] + 1, w[d]);
    }
    double num = 0.0, den = 0.0;
    for (int i = 0; i < L->nx; i++)
        for (int j = 0; j < L->ny; j++) {
            double wij = w[0][i] * w[1][j];
            for (int k = 0; k < L->nz; k++) {
                double ww = wij * w[2][k];
                num += ww * rhs[IX(L, i, j, k)];
                den += ww;
            }
        }
    double c = num / den;
    for (int i = 0; i < L->nx; i++)
        for (int j = 0; j < L->ny; j++)
            for (int k = 0; k < L->nz; k++) rhs[IX(L, i, j, k)] -= c;
    for (int d = 0; d < 3; d++) free(w[d]);
}

/* ------------------------------------------------------------------------------------------ */
/* mgSolve: poisson.cc:100-151                                                                   */
/* ------------------------------------------------------------------------------------------ */


int or_mg_solve(or_mg *m, double *pin, const double *rhs, double *hist, int hist_cap) {
    const or_params *p = &m->p;
    m->vLevel = 0;
    for (int l = 0; l <= p->vcDepth; l++) {
        size_t sz = sizeof(double) * (size_t)m->lv[l].size;
        memset(m->P[l], 0, sz);
        memset(m->R[l], 0, sz);
        memset(m->S[l], 0, sz);
    }
    const level_t *L = &m->lv[0];
    for (int i = 0; i < L->nx; i++)
        for (int j = 0; j < L->ny; j++)
            for (int k = 0; k < L->nz; k++) {
                long c = IX(L, i, j, k);
                m->R[0][c] = rhs[c];
                m->P[0][c] = pin[c];
            }
    if (p->semantics == OR_SEM_PROJECTED && !has_dirichlet(p)) or_mg_project_rhs(m, m->R[0]);
    update_pads_arr(m, m->R[0], 0);
    update_pads_arr(m, m->P[0], 0);
    if (p->semantics == OR_SEM_REFERENCE && !p->testPoisson) {
        double sum = 0.0;
        for (long q = 0; q < L->size; q++) sum += m->R[0][q];
        double mean = sum / (double)L->size;
        for (long q = 0; q < L->size; q++) m->R[0][q] -= mean;
    }
    double prev = 0.0;
    int done = 0;
    for (int it = 0; it < p->vcCount; it++) {
        or_mg_vcycle(m);
        int nt = (p->semantics == OR_SEM_PROJECTED) ? 2 : p->resType;
        double r = or_mg_compute_error(m, nt);
        if (hist && it < hist_cap) hist[it] = r;
        done = it + 1;
        if (p->semantics == OR_SEM_PROJECTED) {
            if (r < p->relTolerance) break;
        } else if (!p->testPoisson) {
            if (fabs(prev - r) < p->mgTolerance) break;
            prev = r;
        }
    }
    memcpy(pin, m->P[0], sizeof(double) * (size_t)L->size);
    return done;
}
