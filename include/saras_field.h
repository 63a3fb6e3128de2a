/*
 * saras_field.h -- C ABI of the field operators of SARAS's lib/field on caller device arrays
 * (gfx950 HIP kernels).
 *
 * These entry points stand for the methods a time step calls on its fields: derivative
 * (lib/field/derivative.h:78-84), vfield (lib/field/vfield.h:76-80), sfield (lib/field/sfield.h:73-76),
 * the boundary classes (the headers of lib/boundary/) with mpidata::syncData, and the implicit Jacobi solves of
 * hydro_d3 / scalar_d3.  Every array is a device pointer the caller owns, in the reference's padded
 * layout: C order, k (z) fastest, logical index -1..n in each direction (lib/field/field.cc:90-91).
 * Core sizes follow the staggering of field.h:70-75 (vfield.cc:64-66):
 *
 *   SR_LOC_NODE  P, T, divergence:  Nx x Ny x Nz nodes
 *   SR_LOC_VX    Vx:  (Nx-1) x Ny x Nz      (x faces, nodes in y and z)
 *   SR_LOC_VY    Vy:  Nx x (Ny-1) x Nz
 *   SR_LOC_VZ    Vz:  Nx x Ny x (Nz-1)
 *
 * with N = 2^ind + 1 nodes per direction, padded by one on each side.
 *
 * Ordering: the operators are enqueued on the handle's stream (sr_fgrid_stream) and return before
 * they finish, except sr_jacobi_visc, which returns with its solve finished (it reads the residual
 * on the host after every sweep, as the reference does).  sr_fgrid_sync waits.  A handle is not
 * re-entrant: calls on one handle must not overlap.  Status codes are those of saras_mg.h.
 */
#ifndef SARAS_FIELD_H
#define SARAS_FIELD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sr_fgrid sr_fgrid;

/* field locations */
#define SR_LOC_NODE 0
#define SR_LOC_VX   1
#define SR_LOC_VY   2
#define SR_LOC_VZ   3

/* the grid of grid::grid (lib/grid/grid.cc:59-136, one sub-domain) with parser's names */
typedef struct sr_fgrid_desc {
    int xInd, yInd, zInd;           /* N = 2^ind + 1 nodes per direction                          */
    double Lx, Ly, Lz;
    int xGrid, yGrid, zGrid;        /* 1 uniform, 2 tanh-stretched (grid.cc:370-422, 442-519)      */
    double betaX, betaY, betaZ;
    int xPer, yPer, zPer;
    int iScheme;                    /* 1: Crank-Nicolson, calcDerivative2* halves the xi^2 part    */
                                    /* (derivative.cc:185-189); otherwise the full expression      */
} sr_fgrid_desc;

/* boundary conditions of one field, walls in the order x-, x+, y-, y+, z- (bottom), z+ (top) */
#define SR_BC_PERIODIC  0
#define SR_BC_DIRICHLET 1
#define SR_BC_NEUMANN   2
#define SR_BC_NONE      3        /* the wall is left as it is: all six SR_BC_NONE is syncData alone */
                                 /* (vfield::syncData, sfield::syncData, plainsf::syncData)          */
typedef struct sr_bc {
    int kind[6];                    /* SR_BC_*; the CC or FC class of the files of lib/boundary/ follows */
                                    /* the field's staggering along the wall's direction           */
    double value[6];                /* Dirichlet wall values (dirichletCC sets the wall slice,     */
                                    /* dirichletFC sets 2 v - F at the image)                      */
} sr_bc;

/* the scalars of the implicit Jacobi solves */
typedef struct sr_jacobi_desc {
    int form;                       /* 0: hydro_d3::solveVx/Vy/Vz's expressions with Re            */
                                    /*    (hydro_d3.cc:370-533); 1: scalar_d3's with a coefficient */
                                    /*    (nu for solveVx/Vy/Vz, kappa for solveT; scalar_d3.cc:373-594) */
    double dt;
    double Re;                      /* form 0                                                      */
    double coef;                    /* form 1                                                      */
    double cnTolerance;             /* stop when max |F - dt coef lap F / 2 - R| over the bulk is  */
                                    /* below it (hydro_d3.cc:396-410)                               */
    int zCollocMetric;              /* 1: the z metric is the collocated xi_z^2 read at node        */
                                    /* indices, as scalar_d3::solveT reads it (scalar_d3.cc:541, 561) */
} sr_jacobi_desc;

/* grid::grid metrics on the device and a stream; derivative::derivative's constant factors
   (derivative.cc:60-105) */
int sr_fgrid_create(const sr_fgrid_desc *desc, sr_fgrid **out);
int sr_fgrid_destroy(sr_fgrid *g);
/* core sizes of a location */
int sr_fgrid_dims(const sr_fgrid *g, int loc, int *nx, int *ny, int *nz);
/* the hipStream_t the operators are enqueued on */
int sr_fgrid_stream(const sr_fgrid *g, void **hip_stream);
int sr_fgrid_sync(sr_fgrid *g);

/* derivative::calcDerivative1_x/_y/_z (derivative.cc:121-164): out = xi_d (F(+1) - F(-1)) / (2 h)
   over the core range along dim and every index (pads included) of the other two directions, with
   the metric of the field's staggering along dim; the pads along dim are set to 0, as on the zeroed
   array every caller passes (vfield.cc:163-253, sfield.cc:113-142) */
int sr_deriv1(sr_fgrid *g, int loc, int dim, const double *F, double *out);
/* derivative::calcDerivative2xx/yy/zz (derivative.cc:178-243): out = xi_dd dF + 0.5 xi_d^2 d2F
   (iScheme 1) over the same range, pads along dim set to 0 */
int sr_deriv2(sr_fgrid *g, int loc, int dim, const double *F, double *out);

/* vfield::computeDiff (vfield.cc:104-146): H_c(core) += d2xx V_c, += d2yy V_c, += d2zz V_c */
int sr_vf_diff(sr_fgrid *g, const double *Vx, const double *Vy, const double *Vz, double *Hx, double *Hy, double *Hz);
/* sfield::computeDiff (sfield.cc:86-100) of a node field */
int sr_sf_diff(sr_fgrid *g, const double *F, double *H);
/* vfield::computeNLin(const vfield &V, plainvf &H) (vfield.cc:163-253): the field U (the vfield the
   method is called on) advected by V: H_c(core) -= (sum of the interpolation slices of V_d) d_d U_c
   / nslices for d = x, y, z; U and V may be the same arrays (hydro_d3.cc:275) */
int sr_vf_nlin(sr_fgrid *g, const double *Ux, const double *Uy, const double *Uz, const double *Vx,
               const double *Vy, const double *Vz, double *Hx, double *Hy, double *Hz);
/* sfield::computeNLin (sfield.cc:113-142): a node field F advected by V */
int sr_sf_nlin(sr_fgrid *g, const double *Vx, const double *Vy, const double *Vz, const double *F, double *H);
/* vfield::divergence (vfield.cc:299-320): div = 0, then its core from the face differences */
int sr_divergence(sr_fgrid *g, const double *Vx, const double *Vy, const double *Vz, double *div);
/* sfield::gradient / plainsf::gradient (sfield.cc:155-171, plainsf.h:92-98): the core of each
   face component, G_c = xi_c(colloc) (F(+1) - F) / h_c; every other entry untouched */
int sr_gradient(sr_fgrid *g, const double *F, double *Gx, double *Gy, double *Gz);

/* boundary conditions of one field as imposeU/V/WBCs and imposeTBCs run them (hydro.cc:386-461,
   scalar.cc:240-255): mpidata::syncData with the rank its own neighbour along periodic x and y
   (mpidata.cc:278-306, faces then edges), the x walls, the y walls, then the top and the bottom
   wall (the files of lib/boundary/).  A periodic direction of the grid takes SR_BC_PERIODIC or SR_BC_NONE on
   both walls, and only a periodic direction takes SR_BC_PERIODIC. */
int sr_bc_fill(sr_fgrid *g, int loc, const sr_bc *bc, double *F);

/* hydro_d3::solveVx/Vy/Vz (hydro_d3.cc:370-533), scalar_d3::solveVx/Vy/Vz/solveT
   (scalar_d3.cc:373-594): Jacobi sweeps of (1 - dt coef lap / 2) F = R over the field's bulk (field.cc:206-219,
   non-periodic node directions drop their end nodes), each followed by the boundary conditions bc,
   until the maximum residual over the bulk is below cnTolerance.  F holds the initial guess and
   the result; G is a caller work array of F's shape (the reference's guess field; its contents on
   return are unspecified).  SR_ERR_NOT_CONVERGED after (Nx-1)(Ny-1)(Nz-1) sweeps without
   convergence, where the reference aborts (hydro_d3.cc:415-421).  *iters = sweeps. */
int sr_jacobi_visc(sr_fgrid *g, int loc, const sr_jacobi_desc *jd, const sr_bc *bc, const double *R, double *F,
                   double *G, int *iters);

#ifdef __cplusplus
}
#endif

#endif
