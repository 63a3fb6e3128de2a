/*
 * CPU ORACLE -- test infrastructure only.
 *
 * A plain-C restatement of the SARAS multigrid Poisson solver (lib/poisson/poisson.cc,
 * lib/poisson/poisson_d3.cc), its grid metrics (lib/grid/grid.cc) and, in field_oracle.c,
 * the field stencils and time-step pieces the solver drivers call each step.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this
 * library, and there only as the checker / CPU baseline.  The product (saras_amd) never
 * links or calls it.
 *
 * Layout: every 3D array is the reference's padded Blitz layout -- shape (Nx+2)(Ny+2)(Nz+2),
 * C order with k (z) fastest, logical index -1..N per dimension (ghost at -1 and N).
 * Every 1D metric array is padded the same way (length N+2, element 0 is index -1).
 */
#ifndef SARAS_ORACLE_H
#define SARAS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* smoothing orders */
#define OR_SMOOTH_LEX    0   /* lexicographic i->j->k Gauss-Seidel, exactly as poisson_d3.cc:137-150 */
#define OR_SMOOTH_RB     1   /* red-black Gauss-Seidel, the ordering of the GPU path                 */
#define OR_SMOOTH_JACOBI 2   /* Jacobi with hierarchy swap, poisson_d3.cc:152-168 (gsSmooth = 0)     */
#define OR_SMOOTH_JACOBI_RB 3 /* Jacobi smoothing with a red-black coarsest solve (GPU, gsSmooth = 0) */

/* solve semantics */
#define OR_SEM_REFERENCE 0   /* mean over the padded RHS, stop on |r_prev - r| < mgTolerance  */
#define OR_SEM_PROJECTED 1   /* RHS projected onto range(L), stop on relative residual       */

typedef struct or_params {
    int xInd, yInd, zInd;      /* N_d = 2^ind + 1 nodes                                       */
    int vcDepth, vcCount;
    int preSmooth, postSmooth;
    int resType;               /* 0, 1, 2 as computeError(normOrder)                            */
    int smoother;              /* OR_SMOOTH_*                                                   */
    int xPer, yPer, zPer;
    int testPoisson;           /* Dirichlet ghosts from wall planes; no mean removal, no early stop */
    int semantics;             /* OR_SEM_*                                                      */
    double mgTolerance;        /* coarsest-solve absolute tolerance and stagnation tolerance   */
    double relTolerance;       /* OR_SEM_PROJECTED: stop when ||f-Lu||/||f|| < relTolerance     */
    double dXi, dEt, dZt;      /* transformed-plane spacing of the finest level               */
    int padMode;               /* residual pads read by restriction: 0 = reference (zero except the
                                  periodic x/y exchange), 1 = consistent with the smoother's ghosts */
    int coarseProject;         /* 1: project the coarsest right-hand side onto the range of its operator
                                  before the coarsest solve (singular problems only)              */
} or_params;

typedef struct or_mg or_mg;

/* ---------- grid (lib/grid/grid.cc) ---------- */
/* Node ("staggered") coordinates and metrics of one direction, padded length n+2.
 * grid_type 1 = uniform, 2 = tanh; beta used only for tanh.                                     */
void or_grid_staggr(int n, double len, int grid_type, double beta,
                    double *x, double *xi_x, double *xixx, double *xix2);
/* Face ("collocated") coordinates and metrics, padded length (n-1)+2.                          */
void or_grid_colloc(int n, double len, int grid_type, double beta,
                    double *x, double *xi_x, double *xixx, double *xix2);

/* ---------- multigrid (lib/poisson) ---------- */
/* metrics: finest-level staggered metric arrays, padded length N+2 each.
 * walls (testPoisson only, else NULL): 6 planes lo/hi for x,y,z, each padded 2D over the other
 * two dims in (a,b) order, e.g. xWall has shape (Ny+2)(Nz+2).                                    */
or_mg *or_mg_create(const or_params *p,
                    const double *xix2, const double *xixx,
                    const double *ety2, const double *etyy,
                    const double *ztz2, const double *ztzz,
                    const double *xWallLo, const double *xWallHi,
                    const double *yWallLo, const double *yWallHi,
                    const double *zWallLo, const double *zWallHi);
void or_mg_destroy(or_mg *m);

/* level data access: which 0 = pressureData, 1 = residualData, 2 = tmpDataArray, 3 = smoothedPres */
double *or_mg_array(or_mg *m, int which, int level);
void    or_mg_dims(const or_mg *m, int level, int *nx, int *ny, int *nz);
void    or_mg_set_level(or_mg *m, int level);
int     or_mg_get_level(const or_mg *m);
void    or_mg_set_zero_bc(or_mg *m, int zeroBC);

/* protected virtuals of multigrid_d3, acting on the current vLevel */
void   or_mg_compute_residual(or_mg *m);
void   or_mg_smooth(or_mg *m, int count);
void   or_mg_solve_coarsest(or_mg *m);
void   or_mg_coarsen(or_mg *m);
void   or_mg_prolong(or_mg *m);
double or_mg_compute_error(or_mg *m, int normOrder);
void   or_mg_impose_bc(or_mg *m);
void   or_mg_update_pads(or_mg *m, int which);
void   or_mg_vcycle(or_mg *m);

/* mgSolve: p (in/out) and rhs are padded level-0 arrays.  Returns the number of V-cycles run;
 * the residual after each cycle goes into hist[0..ret-1] when hist is non-NULL (cap hist_cap).  */
int or_mg_solve(or_mg *m, double *p, const double *rhs, double *hist, int hist_cap);

/* Left null vector of the 1D operator of one direction at the finest level (core length n).
 * Periodic: trapezoid weights; Neumann: solution of L^T w = 0 for the mirrored stencil.          */
void or_null_vector_1d(int n, int periodic, double h, const double *d2, const double *d1, double *w);

/* Project the core of a padded level-0 RHS onto range(L): rhs -= (w.rhs)/(w.1) (tensor w).      */
void or_mg_project_rhs(const or_mg *m, double *rhs);

/* Fine-to-coarse sampled metric of a level (debug / tests): which 0..5 = xix2,xixx,ety2,etyy,ztz2,ztzz */
const double *or_mg_metric(const or_mg *m, int which, int level);
double or_mg_coef(const or_mg *m, int which, int level);  /* 0 hx 1 hy 2 hz 3 hxhy 4 hyhz 5 hzhx 6 hxhyhz */
int    or_mg_max_count(const or_mg *m);

/* ---------- hydro time step (field_oracle.c) ---------- */
typedef struct or_ts_params {
    int xInd, yInd, zInd;
    double Lx, Ly, Lz;
    int xGrid, yGrid, zGrid;
    double betaX, betaY, betaZ;
    int xPer, yPer, zPer;
    int probType;
    double Re, dt, cnTolerance;
    int vcDepth, vcCount, preSmooth, postSmooth, resType, smoother;
    double mgTolerance;
    int scalar;            /* 1: scalar_d3 (temperature with buoyancy forcing, forceType 3)   */
    int rbcType;           /* non-dimensionalisation of scalar.cc:69-86                         */
    double Ra, Pr;
    int semantics;         /* multigrid stop rule of the pressure solve (OR_SEM_*)              */
    double relTolerance;
} or_ts_params;

typedef struct or_ts or_ts;

or_ts *or_ts_create(const or_ts_params *p);
void   or_ts_destroy(or_ts *t);
/* which: 0..2 Vx Vy Vz, 3 P, 4 Pp, 5 mgRHS, 6..8 nseRHS, 9..11 pressure gradient, 12 divergence,
   13 T, 14 tmpRHS */
double *or_ts_array(or_ts *t, int which);
void   or_ts_shape(const or_ts *t, int which, int *nx, int *ny, int *nz);
/* dim 0..2, node 1 = staggered / 0 = collocated, quantity 0 x, 1 xi_x, 2 xi_xx, 3 xi_x^2 */
const double *or_ts_metric(const or_ts *t, int dim, int node, int quantity);
void   or_ts_deriv(or_ts *t, int which, int order, int dim, double *out);
void   or_ts_compute_diff(or_ts *t);
void   or_ts_compute_nlin(or_ts *t);
void   or_ts_gradient(or_ts *t, int from_pp);
void   or_ts_divergence(or_ts *t);
void   or_ts_impose_vbc(or_ts *t, int component);
void   or_ts_init_tgv(or_ts *t);
double or_ts_energy(const or_ts *t);
double or_ts_max_div(or_ts *t);
int    or_ts_solve_v(or_ts *t, int component);
void   or_ts_build_rhs(or_ts *t);
int    or_ts_time_advance(or_ts *t);
int    or_ts_jacobi_iters(const or_ts *t, int component);   /* component 3: temperature */
/* scalar_d3: V = 0, P = 1, T = amp (2u - 1) on the core (u uniform from SplitMix64(seed) in C
   order), then the velocity and temperature BCs */
void   or_ts_init_rbc(or_ts *t, unsigned long long seed, double amp);
void   or_ts_impose_tbc(or_ts *t);
int    or_ts_solve_t(or_ts *t);
void   or_ts_build_trhs(or_ts *t);
double or_ts_nusselt(const or_ts *t);
int    or_ts_mg_cycles(const or_ts *t);
/* field operators on caller arrays of this grid; loc 0 node (P, T), 1..3 the faces of Vx..Vz */
void   or_fd_deriv(or_ts *t, int loc, int order, int dim, const double *F, double *out);
void   or_fd_diff(or_ts *t, int loc, const double *F, double *H);
void   or_fd_nlin(or_ts *t, int loc, const double *F, const double *const *V, double *H);
void   or_fd_gradient(or_ts *t, const double *F, double *G0, double *G1, double *G2);
void   or_fd_divergence(or_ts *t, const double *V0, const double *V1, const double *V2, double *div);
/* kinds per wall (x-, x+, y-, y+, z-, z+): 0 periodic, 1 Dirichlet, 2 Neumann, 3 left as it is */
void   or_fd_bc(or_ts *t, int loc, const int *kinds, const double *values, double *F);

#ifdef __cplusplus
}
#endif
#endif
