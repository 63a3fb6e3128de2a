/*
 * CPU ORACLE -- test infrastructure only (see oracle.h).
 *
 * Restates the field stencils and the hydro time step of SARAS:
 *   lib/field/derivative.cc:121-243   calcDerivative1_{x,y,z}, calcDerivative2{xx,yy,zz}
 *   lib/field/vfield.cc:104-146        computeDiff
 *   lib/field/vfield.cc:163-253        computeNLin (interpolation slices of field.cc:309-575)
 *   lib/field/vfield.cc:299-320        divergence
 *   lib/field/sfield.cc:86-142         computeDiff, computeNLin of a node field
 *   lib/field/sfield.cc:155-171        gradient (plainsf::gradient is the same expression)
 *   lib/parallel/mpidata.cc:82-306     syncData with faceRanks = self along periodic x / y
 *   lib/boundary/*.cc                  periodicCC/FC, dirichletCC/FC, neumannCC/FC
 *   src/solvers/hydro/hydro.cc:287-461 initVBC (probType 2) and imposeU/V/WBCs
 *   src/solvers/hydro/hydro_d3.cc:254-533 timeAdvance, solveVx/Vy/Vz
 *   lib/initial/taylorGreen.cc:54-100  Taylor-Green initial condition
 *   lib/io/tseries.cc:62-178           energy and maximum divergence
 *
 * Staggering (field.h): P is on nodes in every direction (N nodes, "staggered" in the reference's
 * words); Vx sits on x faces (N-1 in x) and on nodes in y, z; Vy and Vz likewise.  Every array is
 * padded by one on each side, C order, k fastest, index -1 first.
 *
 * Derivative alignment: the reference assigns Blitz stencil expressions over the padded field to
 * the core range along the differentiated direction; this restatement takes the centred reading
 * out(r) = metric(r) (F(r+1) - F(r-1)) / (2h) for r in the core range and every other index.
 * One-dimensional metric slices inside index expressions (gradient, divergence) are read at the
 * node or face the derivative is evaluated at.  Both readings are fixed by the closed-form
 * Taylor-Green checks of the reference's tests on uniform grids; on stretched grids they are not
 * pinned by any reference output (parity unpinned there).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    int n[3];      /* core sizes */
    int s[3];      /* 1 = node ("staggered") direction, 0 = face ("collocated") */
} fshape;

static inline long FI(const fshape *f, int i, int j, int k) {
    return ((long)(i + 1) * (f->n[1] + 2) + (j + 1)) * (f->n[2] + 2) + (k + 1);
}
static inline long fsize(const fshape *f) { return (long)(f->n[0] + 2) * (f->n[1] + 2) * (f->n[2] + 2); }

struct or_ts {
    or_ts_params p;
    int N[3];
    double h[3];
    /* metrics: [dim][0 = node (staggr), 1 = face (colloc)][quantity 0 x, 1 xi_x, 2 xixx, 3 xix2] */
    double *met[3][2][4];
    fshape sP, sV[3];
    double *P, *Pp, *mgRHS, *divV;
    double *V[3], *H[3], *G[3], *guess[3], *lap[3], *tmpA, *tmpB, *inter, *deriv;
    or_mg *mg;
    int jacobi_iters[4];
    int mg_cycles;
    /* scalar_d3 */
    double *T, *Trhs, *Tg, *Tlap;
    double nu, kappa, Fb;
};

/* ------------------------------------------------------------------------------------------ */
static double met(const or_ts *t, int d, int node, int q, int i) { return t->met[d][node ? 0 : 1][q][i + 1]; }

or_ts *or_ts_create(const or_ts_params *p) {
    or_ts *t = (or_ts *)calloc(1, sizeof(or_ts));
    t->p = *p;
    int ind[3] = {p->xInd, p->yInd, p->zInd};
    double L[3] = {p->Lx, p->Ly, p->Lz};
    int gt[3] = {p->xGrid, p->yGrid, p->zGrid};
    double beta[3] = {p->betaX, p->betaY, p->betaZ};
    for (int d = 0; d < 3; d++) {
        int n = (1 << ind[d]) + 1;
        t->N[d] = n;
        t->h[d] = 1.0 / (double)(n - 1);
        for (int q = 0; q < 4; q++) {
            t->met[d][0][q] = (double *)calloc((size_t)n + 2, sizeof(double));
            t->met[d][1][q] = (double *)calloc((size_t)n + 1, sizeof(double));
        }
        or_grid_staggr(n, L[d], gt[d], beta[d], t->met[d][0][0], t->met[d][0][1], t->met[d][0][2], t->met[d][0][3]);
        or_grid_colloc(n, L[d], gt[d], beta[d], t->met[d][1][0], t->met[d][1][1], t->met[d][1][2], t->met[d][1][3]);
    }
    for (int d = 0; d < 3; d++) { t->sP.n[d] = t->N[d]; t->sP.s[d] = 1; }
    for (int c = 0; c < 3; c++)
        for (int d = 0; d < 3; d++) {
            t->sV[c].n[d] = (c == d) ? t->N[d] - 1 : t->N[d];
            t->sV[c].s[d] = (c == d) ? 0 : 1;
        }
    long np = fsize(&t->sP);
    t->P = (double *)calloc((size_t)np, sizeof(double));
    t->Pp = (double *)calloc((size_t)np, sizeof(double));
    t->mgRHS = (double *)calloc((size_t)np, sizeof(double));
    t->divV = (double *)calloc((size_t)np, sizeof(double));
    long nmax = 0;
    for (int c = 0; c < 3; c++) {
        long nv = fsize(&t->sV[c]);
        if (nv > nmax) nmax = nv;
        t->V[c] = (double *)calloc((size_t)nv, sizeof(double));
        t->H[c] = (double *)calloc((size_t)nv, sizeof(double));
        t->G[c] = (double *)calloc((size_t)nv, sizeof(double));
        t->guess[c] = (double *)calloc((size_t)nv, sizeof(double));
        t->lap[c] = (double *)calloc((size_t)nv, sizeof(double));
    }
    if (np > nmax) nmax = np;
    t->tmpA = (double *)calloc((size_t)nmax, sizeof(double));
    t->tmpB = (double *)calloc((size_t)nmax, sizeof(double));
    t->inter = (double *)calloc((size_t)nmax, sizeof(double));
    t->deriv = (double *)calloc((size_t)nmax, sizeof(double));
    t->T = (double *)calloc((size_t)np, sizeof(double));
    t->Trhs = (double *)calloc((size_t)np, sizeof(double));
    t->Tg = (double *)calloc((size_t)np, sizeof(double));
    t->Tlap = (double *)calloc((size_t)np, sizeof(double));
    /* scalar.cc:69-86 and buoyantForce.cc:45-55 */
    t->nu = 1.0 / p->Re;
    t->kappa = 0.0;
    t->Fb = 0.0;
    if (p->scalar) {
        switch (p->rbcType) {
            case 1: t->nu = p->Pr; t->kappa = 1.0; t->Fb = p->Ra * p->Pr; break;
            case 3: t->nu = 1.0; t->kappa = 1.0 / p->Pr; t->Fb = p->Ra; break;
            case 4: t->nu = sqrt(p->Pr / p->Ra); t->kappa = 1.0 / sqrt(p->Pr * p->Ra); t->Fb = p->Pr; break;
            default: t->nu = sqrt(p->Pr / p->Ra); t->kappa = 1.0 / sqrt(p->Pr * p->Ra); t->Fb = 1.0; break;
        }
    }
    /* pressure solver on the node grid (hydro_d3.cc:63-66) */
    or_params mp;
    memset(&mp, 0, sizeof(mp));
    mp.xInd = p->xInd; mp.yInd = p->yInd; mp.zInd = p->zInd;
    mp.vcDepth = p->vcDepth; mp.vcCount = p->vcCount;
    mp.preSmooth = p->preSmooth; mp.postSmooth = p->postSmooth;
    mp.resType = p->resType; mp.smoother = p->smoother;
    mp.xPer = p->xPer; mp.yPer = p->yPer; mp.zPer = p->zPer;
    mp.semantics = p->semantics;
    mp.mgTolerance = p->mgTolerance;
    mp.relTolerance = p->relTolerance;
    mp.dXi = t->h[0]; mp.dEt = t->h[1]; mp.dZt = t->h[2];
    t->mg = or_mg_create(&mp, t->met[0][0][3], t->met[0][0][2], t->met[1][0][3], t->met[1][0][2],
                         t->met[2][0][3], t->met[2][0][2], NULL, NULL, NULL, NULL, NULL, NULL);
    return t;
}

void or_ts_destroy(or_ts *t) {
    if (!t) return;
    for (int d = 0; d < 3; d++)
        for (int q = 0; q < 4; q++) { free(t->met[d][0][q]); free(t->met[d][1][q]); }
    free(t->P); free(t->Pp); free(t->mgRHS); free(t->divV);
    for (int c = 0; c < 3; c++) { free(t->V[c]); free(t->H[c]); free(t->G[c]); free(t->guess[c]); free(t->lap[c]); }
    free(t->tmpA); free(t->tmpB); free(t->inter); free(t->deriv);
    free(t->T); free(t->Trhs); free(t->Tg); free(t->Tlap);
    or_mg_destroy(t->mg);
    free(t);
}

double *or_ts_array(or_ts *t, int which) {
    switch (which) {
        case 0: case 1: case 2: return t->V[which];
        case 3: return t->P;
        case 4: return t->Pp;
        case 5: return t->mgRHS;
        case 6: case 7: case 8: return t->H[which - 6];
        case 9: case 10: case 11: return t->G[which - 9];
        case 12: return t->divV;
        case 13: return t->T;
        case 14: return t->Trhs;
        default: return NULL;
    }
}

void or_ts_shape(const or_ts *t, int which, int *nx, int *ny, int *nz) {
    const fshape *f = (which < 3) ? &t->sV[which] : (which >= 6 && which < 12) ? &t->sV[(which - 6) % 3] : &t->sP;
    *nx = f->n[0]; *ny = f->n[1]; *nz = f->n[2];
}

const double *or_ts_metric(const or_ts *t, int dim, int node, int quantity) { return t->met[dim][node ? 0 : 1][quantity]; }
int or_ts_jacobi_iters(const or_ts *t, int c) { return t->jacobi_iters[c]; }
int or_ts_mg_cycles(const or_ts *t) { return t->mg_cycles; }

/* ------------------------------------------------------------------------------------------ */
/* derivatives (derivative.cc:121-243), one field F of shape f into out (same shape)             */
/* ------------------------------------------------------------------------------------------ */
static void d1(const or_ts *t, const fshape *f, const double *F, int d, double *out) {
    long stride[3] = {(long)(f->n[1] + 2) * (f->n[2] + 2), (long)(f->n[2] + 2), 1};
    double ih = 1.0 / t->h[d];
    int lo[3] = {-1, -1, -1}, hi[3] = {f->n[0], f->n[1], f->n[2]};
    lo[d] = 0;
    hi[d] = f->n[d] - 1;
    for (int i = lo[0]; i <= hi[0]; i++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int k = lo[2]; k <= hi[2]; k++) {
                long c = FI(f, i, j, k);
                int r = d == 0 ? i : d == 1 ? j : k;
                double v = (F[c + stride[d]] - F[c - stride[d]]) / 2.0;
                v *= ih;
                out[c] = met(t, d, f->s[d], 1, r) * v;
            }
}

static void d2(const or_ts *t, const fshape *f, const double *F, int d, double *out) {
    long stride[3] = {(long)(f->n[1] + 2) * (f->n[2] + 2), (long)(f->n[2] + 2), 1};
    double ih = 1.0 / t->h[d];
    double ih2 = pow(ih, 2.0);
    int lo[3] = {-1, -1, -1}, hi[3] = {f->n[0], f->n[1], f->n[2]};
    lo[d] = 0;
    hi[d] = f->n[d] - 1;
    for (int i = lo[0]; i <= hi[0]; i++)
        for (int j = lo[1]; j <= hi[1]; j++)
            for (int k = lo[2]; k <= hi[2]; k++) {
                long c = FI(f, i, j, k);
                int r = d == 0 ? i : d == 1 ? j : k;
                double tmp = (F[c + stride[d]] - F[c - stride[d]]) / 2.0 * ih;
                double sec = (F[c + stride[d]] - 2.0 * F[c] + F[c - stride[d]]) * ih2;
                double xx = met(t, d, f->s[d], 2, r), x2 = met(t, d, f->s[d], 3, r);
                out[c] = xx * tmp + 0.5 * x2 * sec;     /* iScheme = 1 (Crank-Nicolson half) */
            }
}

void or_ts_deriv(or_ts *t, int which, int order, int dim, double *out) {
    const fshape *f = (which < 3) ? &t->sV[which] : &t->sP;
    const double *F = (which < 3) ? t->V[which] : t->P;
    memset(out, 0, sizeof(double) * (size_t)fsize(f));
    if (order == 1) d1(t, f, F, dim, out);
    else d2(t, f, F, dim, out);
}

static const fshape *loc_shape(const or_ts *t, int loc) { return loc == 0 ? &t->sP : &t->sV[loc - 1]; }

/* computeDiff of one field at location loc (0 node, 1..3 the faces of Vx..Vz):
   H(core) += d2xx; H(core) += d2yy; H(core) += d2zz, each into a zeroed temporary first
   (vfield.cc:104-146 per component, sfield.cc:86-100) */
void or_fd_diff(or_ts *t, int loc, const double *F, double *H) {
    const fshape *f = loc_shape(t, loc);
    for (int d = 0; d < 3; d++) {
        memset(t->deriv, 0, sizeof(double) * (size_t)fsize(f));
        d2(t, f, F, d, t->deriv);
        for (int i = 0; i < f->n[0]; i++)
            for (int j = 0; j < f->n[1]; j++)
                for (int k = 0; k < f->n[2]; k++) H[FI(f, i, j, k)] += t->deriv[FI(f, i, j, k)];
    }
}

/* vfield::computeDiff (vfield.cc:104-146): H(core) += d2xx + d2yy + d2zz, component by component */
void or_ts_compute_diff(or_ts *t) {
    for (int c = 0; c < 3; c++) or_fd_diff(t, c + 1, t->V[c], t->H[c]);
}

/* interpolation slices of velocity component d onto the location of component c (field.cc):
   d == c: the core itself; otherwise {core, d-1, c+1, (c+1, d-1)} */
static int slices(int c, int d, int off[4][3]) {
    memset(off, 0, sizeof(int) * 12);
    if (c == d) return 1;
    off[1][d] = -1;
    off[2][c] = 1;
    off[3][c] = 1;
    off[3][d] = -1;
    return 4;
}

/* computeNLin of the field F at location loc advected by V (vfield.cc:163-253 for a velocity
   component, sfield.cc:113-142 for a node field): for d = x, y, z in turn,
   H(core) -= (sum of the interpolation slices of V_d) d_d(F) / nslices.  Slices of V_d onto the
   location of component c are those of slices(); onto the nodes, the two faces on either side
   (field.cc:412-434): the core index, then one back along d. */
void or_fd_nlin(or_ts *t, int loc, const double *F, const double *const *V, double *H) {
    const fshape *f = loc_shape(t, loc);
    for (int d = 0; d < 3; d++) {
        int off[4][3];
        int ns;
        if (loc == 0) {
            memset(off, 0, sizeof(off));
            off[1][d] = -1;
            ns = 2;
        } else {
            ns = slices(loc - 1, d, off);
        }
        const fshape *fd = &t->sV[d];
        memset(t->inter, 0, sizeof(double) * (size_t)fsize(f));
        for (int s = 0; s < ns; s++)
            for (int i = 0; i < f->n[0]; i++)
                for (int j = 0; j < f->n[1]; j++)
                    for (int k = 0; k < f->n[2]; k++)
                        t->inter[FI(f, i, j, k)] += V[d][FI(fd, i + off[s][0], j + off[s][1], k + off[s][2])];
        memset(t->deriv, 0, sizeof(double) * (size_t)fsize(f));
        d1(t, f, F, d, t->deriv);
        for (int i = 0; i < f->n[0]; i++)
            for (int j = 0; j < f->n[1]; j++)
                for (int k = 0; k < f->n[2]; k++) {
                    long q = FI(f, i, j, k);
                    H[q] -= t->inter[q] * t->deriv[q] / (double)ns;
                }
    }
}

/* vfield::computeNLin (vfield.cc:163-253) of the velocity by itself: H_c(core) -= (sum of slices
   of V_d) d_d(V_c) / nslices */
void or_ts_compute_nlin(or_ts *t) {
    const double *V[3] = {t->V[0], t->V[1], t->V[2]};
    for (int c = 0; c < 3; c++) or_fd_nlin(t, c + 1, t->V[c], V, t->H[c]);
}

/* sfield::gradient (sfield.cc:155-171): grad_c on the core of V_c = metric (P(+1) - P) / h */
static void gradient(or_ts *t, const double *Pf, double **G) {
    for (int c = 0; c < 3; c++) {
        const fshape *f = &t->sV[c];
        long sp[3] = {(long)(t->sP.n[1] + 2) * (t->sP.n[2] + 2), (long)(t->sP.n[2] + 2), 1};
        for (int i = 0; i < f->n[0]; i++)
            for (int j = 0; j < f->n[1]; j++)
                for (int k = 0; k < f->n[2]; k++) {
                    int r = c == 0 ? i : c == 1 ? j : k;
                    long q = FI(&t->sP, i, j, k);
                    G[c][FI(f, i, j, k)] = met(t, c, 0, 1, r) * (Pf[q + sp[c]] - Pf[q]) / t->h[c];
                }
    }
}

void or_ts_gradient(or_ts *t, int from_pp) {
    for (int c = 0; c < 3; c++) memset(t->G[c], 0, sizeof(double) * (size_t)fsize(&t->sV[c]));
    gradient(t, from_pp ? t->Pp : t->P, t->G);
}

/* vfield::divergence (vfield.cc:299-320) into divV (zeroed first) */
void or_ts_divergence(or_ts *t) {
    const fshape *p = &t->sP;
    memset(t->divV, 0, sizeof(double) * (size_t)fsize(p));
    for (int i = 0; i < p->n[0]; i++)
        for (int j = 0; j < p->n[1]; j++)
            for (int k = 0; k < p->n[2]; k++) {
                double a = met(t, 0, 1, 1, i) * (t->V[0][FI(&t->sV[0], i, j, k)] - t->V[0][FI(&t->sV[0], i - 1, j, k)]) / t->h[0];
                double b = met(t, 1, 1, 1, j) * (t->V[1][FI(&t->sV[1], i, j, k)] - t->V[1][FI(&t->sV[1], i, j - 1, k)]) / t->h[1];
                double cz = met(t, 2, 1, 1, k) * (t->V[2][FI(&t->sV[2], i, j, k)] - t->V[2][FI(&t->sV[2], i, j, k - 1)]) / t->h[2];
                t->divV[FI(p, i, j, k)] = a + b + cz;
            }
}

/* sfield::gradient on caller arrays (sfield.cc:155-171): the core of each component only */
void or_fd_gradient(or_ts *t, const double *F, double *G0, double *G1, double *G2) {
    double *G[3] = {G0, G1, G2};
    gradient(t, F, G);
}

/* vfield::divergence on caller arrays (vfield.cc:299-320): divV = 0, then its core */
void or_fd_divergence(or_ts *t, const double *V0, const double *V1, const double *V2, double *div) {
    const fshape *p = &t->sP;
    const double *V[3] = {V0, V1, V2};
    memset(div, 0, sizeof(double) * (size_t)fsize(p));
    for (int i = 0; i < p->n[0]; i++)
        for (int j = 0; j < p->n[1]; j++)
            for (int k = 0; k < p->n[2]; k++) {
                double a = met(t, 0, 1, 1, i) * (V[0][FI(&t->sV[0], i, j, k)] - V[0][FI(&t->sV[0], i - 1, j, k)]) / t->h[0];
                double b = met(t, 1, 1, 1, j) * (V[1][FI(&t->sV[1], i, j, k)] - V[1][FI(&t->sV[1], i, j - 1, k)]) / t->h[1];
                double cz = met(t, 2, 1, 1, k) * (V[2][FI(&t->sV[2], i, j, k)] - V[2][FI(&t->sV[2], i, j, k - 1)]) / t->h[2];
                div[FI(p, i, j, k)] = a + b + cz;
            }
}

/* derivative::calcDerivative1_* / calcDerivative2* of a caller field at location loc into out,
   zeroed first as every caller zeroes it (vfield.cc:104-253, sfield.cc:86-142) */
void or_fd_deriv(or_ts *t, int loc, int order, int dim, const double *F, double *out) {
    const fshape *f = loc_shape(t, loc);
    memset(out, 0, sizeof(double) * (size_t)fsize(f));
    if (order == 1) d1(t, f, F, dim, out);
    else d2(t, f, F, dim, out);
}

/* ------------------------------------------------------------------------------------------ */
/* ghost layers                                                                                  */
/* ------------------------------------------------------------------------------------------ */
/* mpidata::syncData (mpidata.cc:278-306) with faceRanks / edgeRanks = self along periodic x and y:
   faces over the core of the other two directions, then the x-y edges over the core z range */
static void sync_data(const or_ts *t, const fshape *f, double *F) {
    int nx = f->n[0], ny = f->n[1], nz = f->n[2];
    int sx = f->s[0], sy = f->s[1];
    int xp = t->p.xPer, yp = t->p.yPer;
    if (xp)
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < nz; k++) {
                F[FI(f, -1, j, k)] = F[FI(f, nx - 1 - sx, j, k)];
                F[FI(f, nx, j, k)] = F[FI(f, sx, j, k)];
            }
    if (yp)
        for (int i = 0; i < nx; i++)
            for (int k = 0; k < nz; k++) {
                F[FI(f, i, -1, k)] = F[FI(f, i, ny - 1 - sy, k)];
                F[FI(f, i, ny, k)] = F[FI(f, i, sy, k)];
            }
    /* edgeRanks are never cleared by checkPeriodic (hydro.cc:174-212): on one rank the edge
       exchange always runs, whatever the periodicity */
    for (int k = 0; k < nz; k++) {
            F[FI(f, -1, -1, k)] = F[FI(f, nx - 1 - sx, ny - 1 - sy, k)];
            F[FI(f, -1, ny, k)] = F[FI(f, nx - 1 - sx, sy, k)];
            F[FI(f, nx, -1, k)] = F[FI(f, sx, ny - 1 - sy, k)];
            F[FI(f, nx, ny, k)] = F[FI(f, sx, sy, k)];
        }
}

/* Boundary of one wall (0..5) of field f: the wall slice is the layer just outside the bulk
   (field.cc:180-290).  kind: 0 periodicCC, 1 periodicFC, 2 dirichletCC, 3 dirichletFC,
   4 neumannCC, 5 neumannFC (lib/boundary/*.cc). */
static void wall_bc(const fshape *f, int per, double *F, int wall, int kind, double value) {
    int d = wall / 2, hi = wall & 1;
    int n = f->n[d];
    /* bulk along d (field.cc:206-219): node directions that are not periodic drop their end nodes */
    int blo = 0, bhi = n - 1;
    if (f->s[d] && !per) { blo = 1; bhi = n - 2; }
    int w = hi ? bhi + 1 : blo - 1;           /* wall index */
    int wo = hi ? blo - 1 : bhi + 1;          /* the opposite wall */
    int src;
    switch (kind) {
        case 0: src = hi ? wo + 2 : wo - 2; break;        /* periodicCC: second point from the other wall */
        case 1: src = hi ? wo + 1 : wo - 1; break;        /* periodicFC */
        case 3: case 4: case 5: src = hi ? w - 1 : w + 1; break;
        default: src = w; break;
    }
    int lo[3] = {-1, -1, -1}, up[3] = {f->n[0], f->n[1], f->n[2]};
    lo[d] = up[d] = w;
    for (int i = lo[0]; i <= up[0]; i++)
        for (int j = lo[1]; j <= up[1]; j++)
            for (int k = lo[2]; k <= up[2]; k++) {
                int q[3] = {i, j, k};
                q[d] = src;
                long dst = FI(f, i, j, k), s = FI(f, q[0], q[1], q[2]);
                switch (kind) {
                    case 0: case 1: case 4: case 5: F[dst] = F[s]; break;
                    case 2: F[dst] = value; break;
                    case 3: F[dst] = 2.0 * value - F[s]; break;
                }
            }
}

/* hydro::imposeU/V/WBCs (hydro.cc:386-461) with the probType 2 walls of initVBC (hydro.cc:287-371):
   walls normal to the component are no-penetration (dirichletFC 0), the others no-slip
   (dirichletCC 0); z walls periodicCC / periodicFC when zPer */
void or_ts_impose_vbc(or_ts *t, int c) {
    const fshape *f = &t->sV[c];
    double *F = t->V[c];
    int per[3] = {t->p.xPer, t->p.yPer, t->p.zPer};
    sync_data(t, f, F);
    for (int d = 0; d < 2; d++) {
        if (per[d]) continue;
        int kind = (d == c) ? 3 : 2;
        wall_bc(f, 0, F, 2 * d, kind, 0.0);
        wall_bc(f, 0, F, 2 * d + 1, kind, 0.0);
    }
    /* top then bottom, as imposeUBCs calls them */
    if (per[2]) {
        int kind = (c == 2) ? 1 : 0;
        wall_bc(f, 1, F, 5, kind, 0.0);
        wall_bc(f, 1, F, 4, kind, 0.0);
    } else {
        int kind = (c == 2) ? 3 : 2;
        double top = (t->p.probType == 1 && c == 0) ? 1.0 : 0.0;
        wall_bc(f, 0, F, 5, kind, top);
        wall_bc(f, 0, F, 4, kind, 0.0);
    }
}

/* Boundary conditions of one field at location loc, as a field's imposeBC runs them
   (hydro.cc:386-461, scalar.cc:240-255): syncData, the x walls, the y walls, then the top and the
   bottom wall.  kinds[w] per wall (0 x-, 1 x+, 2 y-, 3 y+, 4 z-, 5 z+): 0 periodic, 1 Dirichlet,
   2 Neumann, 3 left as it is, the CC or FC class chosen by the field's staggering along the wall's direction
   (lib/boundary/*.cc).  Periodic x and y walls are the syncData exchange; periodic z walls are
   periodicCC / periodicFC. */
void or_fd_bc(or_ts *t, int loc, const int *kinds, const double *values, double *F) {
    const fshape *f = loc_shape(t, loc);
    int per[3] = {t->p.xPer, t->p.yPer, t->p.zPer};
    sync_data(t, f, F);
    for (int d = 0; d < 2; d++) {
        if (per[d]) continue;
        for (int w = 2 * d; w <= 2 * d + 1; w++) {
            if (kinds[w] == 3) continue;
            int node = f->s[d];
            int kind = kinds[w] == 1 ? (node ? 2 : 3) : (node ? 4 : 5);
            wall_bc(f, 0, F, w, kind, values[w]);
        }
    }
    int order[2] = {5, 4};
    for (int q = 0; q < 2; q++) {
        int w = order[q];
        if (kinds[w] == 3) continue;
        int node = f->s[2];
        int kind;
        if (per[2]) kind = node ? 0 : 1;
        else kind = kinds[w] == 1 ? (node ? 2 : 3) : (node ? 4 : 5);
        wall_bc(f, per[2], F, w, kind, values[w]);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* initial condition and diagnostics                                                            */
/* ------------------------------------------------------------------------------------------ */
/* hydro_d3 constructor (hydro_d3.cc:86-118): P = 1, Taylor-Green V (taylorGreen.cc:54-100),
   then the velocity BCs */
void or_ts_init_tgv(or_ts *t) {
    long np = fsize(&t->sP);
    for (long q = 0; q < np; q++) { t->P[q] = 1.0; t->Pp[q] = 0.0; }
    double L[3] = {t->p.Lx, t->p.Ly, t->p.Lz};
    for (int c = 0; c < 3; c++) {
        const fshape *f = &t->sV[c];
        for (int i = -1; i <= f->n[0]; i++)
            for (int j = -1; j <= f->n[1]; j++)
                for (int k = -1; k <= f->n[2]; k++) {
                    double x = met(t, 0, f->s[0], 0, i), y = met(t, 1, f->s[1], 0, j), z = met(t, 2, f->s[2], 0, k);
                    double v = 0.0;
                    if (c == 0) v = sin(2.0 * M_PI * x / L[0]) * cos(2.0 * M_PI * y / L[1]) * cos(2.0 * M_PI * z / L[2]);
                    if (c == 1) v = -cos(2.0 * M_PI * x / L[0]) * sin(2.0 * M_PI * y / L[1]) * cos(2.0 * M_PI * z / L[2]);
                    t->V[c][FI(f, i, j, k)] = v;
                }
    }
    for (int c = 0; c < 3; c++) or_ts_impose_vbc(t, c);
}

/* tseries energy (tseries.cc:62-107, 158-169): volume-weighted mean of the sum of squared
   face-to-node averages, over every core node (duplicated periodic nodes included), no 1/2 */
double or_ts_energy(const or_ts *t) {
    double vol = 0.0, en = 0.0;
    const fshape *p = &t->sP;
    for (int i = 0; i < p->n[0]; i++)
        for (int j = 0; j < p->n[1]; j++)
            for (int k = 0; k < p->n[2]; k++) {
                double dv = (t->h[0] / met(t, 0, 0, 1, i)) * (t->h[1] / met(t, 1, 0, 1, j)) * (t->h[2] / met(t, 2, 0, 1, k));
                vol += dv;
            }
    for (int i = 0; i < p->n[0]; i++)
        for (int j = 0; j < p->n[1]; j++)
            for (int k = 0; k < p->n[2]; k++) {
                double u = (t->V[0][FI(&t->sV[0], i - 1, j, k)] + t->V[0][FI(&t->sV[0], i, j, k)]) / 2.0;
                double v = (t->V[1][FI(&t->sV[1], i, j - 1, k)] + t->V[1][FI(&t->sV[1], i, j, k)]) / 2.0;
                double w = (t->V[2][FI(&t->sV[2], i, j, k - 1)] + t->V[2][FI(&t->sV[2], i, j, k)]) / 2.0;
                en += (pow(u, 2.0) + pow(v, 2.0) + pow(w, 2.0)) *
                      (t->h[0] / met(t, 0, 0, 1, i)) * (t->h[1] / met(t, 1, 0, 1, j)) * (t->h[2] / met(t, 2, 0, 1, k));
            }
    return en / vol;
}

/* maximum of the divergence over the node core (tseries.cc:137-145, plainsf::fxMax) */
double or_ts_max_div(or_ts *t) {
    or_ts_divergence(t);
    const fshape *p = &t->sP;
    double m = -1e300;
    for (int i = 0; i < p->n[0]; i++)
        for (int j = 0; j < p->n[1]; j++)
            for (int k = 0; k < p->n[2]; k++)
                if (t->divV[FI(p, i, j, k)] > m) m = t->divV[FI(p, i, j, k)];
    return m;
}

/* ------------------------------------------------------------------------------------------ */
/* implicit Jacobi viscous solve (hydro_d3.cc:370-533)                                          */
/* ------------------------------------------------------------------------------------------ */
int or_ts_solve_v(or_ts *t, int c) {
    const fshape *f = &t->sV[c];
    double *V = t->V[c], *R = t->H[c], *Gs = t->guess[c], *Lp = t->lap[c];
    double hx = t->h[0], hy = t->h[1], hz = t->h[2];
    double hx2hy2 = pow(hx, 2.0) * pow(hy, 2.0), hy2hz2 = pow(hy, 2.0) * pow(hz, 2.0);
    double hz2hx2 = pow(hz, 2.0) * pow(hx, 2.0);
    double hx2hy2hz2 = pow(hx, 2.0) * pow(hy, 2.0) * pow(hz, 2.0);
    double dt = t->p.dt, Re = t->p.Re, nu = t->nu;
    double visc = t->p.scalar ? nu : 1.0 / Re;
    int per[3] = {t->p.xPer, t->p.yPer, t->p.zPer};
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = 0;
        hi[d] = f->n[d] - 1;
        if (f->s[d] && !per[d]) { lo[d] = 1; hi[d] = f->n[d] - 2; }
    }
    long st[3] = {(long)(f->n[1] + 2) * (f->n[2] + 2), (long)(f->n[2] + 2), 1};
    int maxIterations = (t->N[0] - 1) * (t->N[1] - 1) * (t->N[2] - 1);
    int iterCount = 0, sweeps = 0;
    while (1) {
        for (int i = lo[0]; i <= hi[0]; i++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int k = lo[2]; k <= hi[2]; k++) {
                    long q = FI(f, i, j, k);
                    double mx = met(t, 0, f->s[0], 3, i), my = met(t, 1, f->s[1], 3, j), mz = met(t, 2, f->s[2], 3, k);
                    double a = hy2hz2 * mx * (V[q + st[0]] + V[q - st[0]]) + hz2hx2 * my * (V[q + st[1]] + V[q - st[1]]) +
                               hx2hy2 * mz * (V[q + 1] + V[q - 1]);
                    double d = hy2hz2 * mx + hz2hx2 * my + hx2hy2 * mz;
                    if (t->p.scalar)      /* scalar_d3.cc:378-386 */
                        Gs[q] = (a * dt * nu / (hx2hy2hz2 * 2.0) + R[q]) / (1.0 + dt * nu * (d) / hx2hy2hz2);
                    else                  /* hydro_d3.cc:378-386 */
                        Gs[q] = (a * dt / (hx2hy2hz2 * 2.0 * Re) + R[q]) / (1.0 + dt * (d) / (Re * hx2hy2hz2));
                }
        memcpy(V, Gs, sizeof(double) * (size_t)fsize(f));
        or_ts_impose_vbc(t, c);
        double maxError = 0.0;
        for (int i = lo[0]; i <= hi[0]; i++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int k = lo[2]; k <= hi[2]; k++) {
                    long q = FI(f, i, j, k);
                    double mx = met(t, 0, f->s[0], 3, i), my = met(t, 1, f->s[1], 3, j), mz = met(t, 2, f->s[2], 3, k);
                    Lp[q] = V[q] - (mx * (V[q + st[0]] - 2.0 * V[q] + V[q - st[0]]) / (hx * hx) +
                                    my * (V[q + st[1]] - 2.0 * V[q] + V[q - st[1]]) / (hy * hy) +
                                    mz * (V[q + 1] - 2.0 * V[q] + V[q - 1]) / (hz * hz)) * 0.5 * dt * visc;
                    double e = fabs(Lp[q] - R[q]);
                    if (e > maxError) maxError = e;
                }
        sweeps += 1;
        if (maxError < t->p.cnTolerance) break;
        iterCount += 1;
        if (iterCount > maxIterations) return -1;   /* the reference aborts here (hydro_d3.cc:415-421) */
    }
    t->jacobi_iters[c] = sweeps;
    return sweeps;
}

/* ------------------------------------------------------------------------------------------ */
/* hydro_d3::timeAdvance (hydro_d3.cc:254-368), zero forcing                                     */
/* ------------------------------------------------------------------------------------------ */
/* The first half: nseRHS = V + dt (nu-free diffusion / Re - advection - grad P), synced.       */
void or_ts_build_rhs(or_ts *t) {
    /* hydro_d3 multiplies by 1/Re (hydro_d3.cc:259-262), scalar_d3 by nu (scalar_d3.cc:254-256) */
    double visc = t->p.scalar ? t->nu : 1.0 / t->p.Re, dt = t->p.dt;
    for (int c = 0; c < 3; c++) memset(t->H[c], 0, sizeof(double) * (size_t)fsize(&t->sV[c]));
    or_ts_compute_diff(t);
    for (int c = 0; c < 3; c++) {
        long n = fsize(&t->sV[c]);
        for (long q = 0; q < n; q++) t->H[c][q] *= visc;
    }
    or_ts_compute_nlin(t);
    if (t->p.scalar) {
        /* buoyantForce::addForcing (buoyantForce.cc:59-68): Vz += Fb (T(k) + T(k+1)) / 2 */
        const fshape *f = &t->sV[2];
        for (int i = 0; i < f->n[0]; i++)
            for (int j = 0; j < f->n[1]; j++)
                for (int k = 0; k < f->n[2]; k++) {
                    double inter = 0.0;
                    inter += t->T[FI(&t->sP, i, j, k)];
                    inter += t->T[FI(&t->sP, i, j, k + 1)];
                    inter /= 2.0;
                    t->H[2][FI(f, i, j, k)] += t->Fb * inter;
                }
    }
    or_ts_gradient(t, 0);
    for (int c = 0; c < 3; c++) {
        long n = fsize(&t->sV[c]);
        for (long q = 0; q < n; q++) {
            t->H[c][q] -= t->G[c][q];
            t->H[c][q] *= dt;
            t->H[c][q] += t->V[c][q];
        }
        sync_data(t, &t->sV[c], t->H[c]);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* scalar_d3: temperature equation, its BCs and the RBC initial state                          */
/* ------------------------------------------------------------------------------------------ */
/* scalar::imposeTBCs (scalar.cc:240-255) with the probType 5 walls of initTBC (scalar.cc:196-228):
   syncData, adiabatic (neumannCC) x and y walls, then the cold top (0) and the hot bottom (1) */
void or_ts_impose_tbc(or_ts *t) {
    const fshape *f = &t->sP;
    double *F = t->T;
    sync_data(t, f, F);
    if (!t->p.xPer) { wall_bc(f, 0, F, 0, 4, 0.0); wall_bc(f, 0, F, 1, 4, 0.0); }
    if (!t->p.yPer) { wall_bc(f, 0, F, 2, 4, 0.0); wall_bc(f, 0, F, 3, 4, 0.0); }
    if (t->p.zPer) {
        wall_bc(f, 1, F, 5, 0, 0.0);
        wall_bc(f, 1, F, 4, 0, 0.0);
    } else {
        /* probType 5: hot bottom, cold top; probType 6 the reverse (scalar.cc:206-227) */
        double top = t->p.probType == 6 ? 1.0 : 0.0, bottom = t->p.probType == 6 ? 0.0 : 1.0;
        wall_bc(f, 0, F, 5, 2, top);
        wall_bc(f, 0, F, 4, 2, bottom);
    }
}

/* the temperature right-hand side of scalar_d3::timeAdvance (scalar_d3.cc:349-360):
   tmpRHS = T + dt (kappa sum d2 T - sum_d (V_d interpolated) dT/dx_d), then syncData.  The
   interpolation of V_d onto the nodes averages the faces on either side (field.cc:412-434). */
void or_ts_build_trhs(or_ts *t) {
    const fshape *p = &t->sP;
    long np = fsize(p);
    memset(t->Trhs, 0, sizeof(double) * (size_t)np);
    or_fd_diff(t, 0, t->T, t->Trhs);
    for (long q = 0; q < np; q++) t->Trhs[q] *= t->kappa;
    const double *V[3] = {t->V[0], t->V[1], t->V[2]};
    or_fd_nlin(t, 0, t->T, V, t->Trhs);
    for (long q = 0; q < np; q++) {
        t->Trhs[q] *= t->p.dt;
        t->Trhs[q] += t->T[q];
    }
    sync_data(t, p, t->Trhs);
}

/* scalar_d3::solveT (scalar_d3.cc:533-594).  The z metric is ztz2Colloc read at node indices, as
   the reference does (scalar_d3.cc:541, 561). */
int or_ts_solve_t(or_ts *t) {
    const fshape *f = &t->sP;
    double *T = t->T, *R = t->Trhs, *Gs = t->Tg, *Lp = t->Tlap;
    double hx = t->h[0], hy = t->h[1], hz = t->h[2];
    double hx2hy2 = pow(hx, 2.0) * pow(hy, 2.0), hy2hz2 = pow(hy, 2.0) * pow(hz, 2.0);
    double hz2hx2 = pow(hz, 2.0) * pow(hx, 2.0);
    double hx2hy2hz2 = pow(hx, 2.0) * pow(hy, 2.0) * pow(hz, 2.0);
    double dt = t->p.dt, kappa = t->kappa;
    int per[3] = {t->p.xPer, t->p.yPer, t->p.zPer};
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
        lo[d] = 0;
        hi[d] = f->n[d] - 1;
        if (!per[d]) { lo[d] = 1; hi[d] = f->n[d] - 2; }
    }
    long st[3] = {(long)(f->n[1] + 2) * (f->n[2] + 2), (long)(f->n[2] + 2), 1};
    int maxIterations = (t->N[0] - 1) * (t->N[1] - 1) * (t->N[2] - 1);
    int iterCount = 0, sweeps = 0;
    while (1) {
        for (int i = lo[0]; i <= hi[0]; i++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int k = lo[2]; k <= hi[2]; k++) {
                    long q = FI(f, i, j, k);
                    double mx = met(t, 0, 1, 3, i), my = met(t, 1, 1, 3, j), mz = met(t, 2, 0, 3, k);
                    double a = hy2hz2 * mx * (T[q + st[0]] + T[q - st[0]]) + hz2hx2 * my * (T[q + st[1]] + T[q - st[1]]) +
                               hx2hy2 * mz * (T[q + 1] + T[q - 1]);
                    double d = hy2hz2 * mx + hz2hx2 * my + hx2hy2 * mz;
                    Gs[q] = (a * dt * kappa / (hx2hy2hz2 * 2.0) + R[q]) / (1.0 + dt * kappa * (d) / hx2hy2hz2);
                }
        memcpy(T, Gs, sizeof(double) * (size_t)fsize(f));
        or_ts_impose_tbc(t);
        double maxError = 0.0;
        for (int i = lo[0]; i <= hi[0]; i++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int k = lo[2]; k <= hi[2]; k++) {
                    long q = FI(f, i, j, k);
                    double mx = met(t, 0, 1, 3, i), my = met(t, 1, 1, 3, j), mz = met(t, 2, 0, 3, k);
                    Lp[q] = T[q] - (mx * (T[q + st[0]] - 2.0 * T[q] + T[q - st[0]]) / (hx * hx) +
                                    my * (T[q + st[1]] - 2.0 * T[q] + T[q - st[1]]) / (hy * hy) +
                                    mz * (T[q + 1] - 2.0 * T[q] + T[q - 1]) / (hz * hz)) * 0.5 * dt * kappa;
                    double e = fabs(Lp[q] - R[q]);
                    if (e > maxError) maxError = e;
                }
        sweeps += 1;
        if (maxError < t->p.cnTolerance) break;
        iterCount += 1;
        if (iterCount > maxIterations) return -1;
    }
    t->jacobi_iters[3] = sweeps;
    return sweeps;
}

static unsigned long long sm_next(unsigned long long *s) {
    *s += 0x9E3779B97F4A7C15ull;
    unsigned long long z = *s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* scalar_d3 constructor (scalar_d3.cc:85-107): P = 1, T = 0 (here plus a seeded perturbation on
   the core, the convection case's initial state), V = 0, then the velocity and temperature BCs */
void or_ts_init_rbc(or_ts *t, unsigned long long seed, double amp) {
    long np = fsize(&t->sP);
    for (long q = 0; q < np; q++) { t->P[q] = 1.0; t->Pp[q] = 0.0; t->T[q] = 0.0; }
    for (int c = 0; c < 3; c++) memset(t->V[c], 0, sizeof(double) * (size_t)fsize(&t->sV[c]));
    unsigned long long st = seed;
    const fshape *p = &t->sP;
    for (int i = 0; i < p->n[0]; i++)
        for (int j = 0; j < p->n[1]; j++)
            for (int k = 0; k < p->n[2]; k++) {
                double u = ((double)(sm_next(&st) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
                t->T[FI(p, i, j, k)] = amp * (2.0 * u - 1.0);
            }
    for (int c = 0; c < 3; c++) or_ts_impose_vbc(t, c);
    or_ts_impose_tbc(t);
}

/* tseries::writeTSData(T, nu, kappa) (tseries.cc:195-246): Nu = 1 + <w T> / kappa, w averaged
   from the two z faces of each node, volume-weighted over every core node */
double or_ts_nusselt(const or_ts *t) {
    double vol = 0.0, uzt = 0.0;
    const fshape *p = &t->sP;
    for (int i = 0; i < p->n[0]; i++)
        for (int j = 0; j < p->n[1]; j++)
            for (int k = 0; k < p->n[2]; k++) {
                double dv = (t->h[0] / met(t, 0, 0, 1, i)) * (t->h[1] / met(t, 1, 0, 1, j)) * (t->h[2] / met(t, 2, 0, 1, k));
                vol += dv;
                uzt += ((t->V[2][FI(&t->sV[2], i, j, k)] + t->V[2][FI(&t->sV[2], i, j, k - 1)]) / 2.0) *
                       t->T[FI(p, i, j, k)] * dv;
            }
    return 1.0 + (uzt / vol) / t->kappa;
}

int or_ts_time_advance(or_ts *t) {
    double dt = t->p.dt;
    or_ts_build_rhs(t);
    for (int c = 0; c < 3; c++)
        if (or_ts_solve_v(t, c) < 0) return -1;
    or_ts_divergence(t);
    long np = fsize(&t->sP);
    for (long q = 0; q < np; q++) t->mgRHS[q] = t->divV[q] * (1.0 / dt);
    t->mg_cycles = or_mg_solve(t->mg, t->Pp, t->mgRHS, NULL, 0);
    sync_data(t, &t->sP, t->Pp);
    for (long q = 0; q < np; q++) t->P[q] += t->Pp[q];
    or_ts_gradient(t, 1);
    for (int c = 0; c < 3; c++) {
        long n = fsize(&t->sV[c]);
        for (long q = 0; q < n; q++) {
            t->G[c][q] *= dt;
            t->V[c][q] -= t->G[c][q];
        }
    }
    if (t->p.scalar) {
        or_ts_build_trhs(t);
        if (or_ts_solve_t(t) < 0) return -1;
    }
    for (int c = 0; c < 3; c++) or_ts_impose_vbc(t, c);
    if (t->p.scalar) or_ts_impose_tbc(t);
    return 0;
}
