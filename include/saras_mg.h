/*
 * saras_mg.h -- C ABI of the MI355X multigrid Poisson solver (gfx950 HIP kernels).
 *
 * Drop-in boundary for SARAS's pressure solve.  Each entry point names the reference interface it
 * replaces (paths relative to the SARAS source tree).  No C++ or torch types cross this boundary:
 * plain pointers, sizes and int status codes (0 = SR_OK).
 *
 * Array layout at the boundary is the reference's padded Blitz layout: shape (Nx+2)(Ny+2)(Nz+2),
 * C order with k (z) fastest, logical index -1..N in each direction (lib/poisson/poisson.cc:300-301,
 * stagFull).  Internally the solver keeps every level in a red-black split layout; the boundary
 * packs and unpacks.
 *
 * Threading: a handle is not re-entrant (like the reference object, whose vLevel and level
 * hierarchies are member state, lib/poisson/poisson.h:56-73); calls on one handle must not overlap.
 */
#ifndef SARAS_MG_H
#define SARAS_MG_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes; the reference prints and calls exit(0) instead (e.g. lib/grid/grid.cc:65-69) */
#define SR_OK                0
#define SR_ERR_ARG           1   /* invalid argument or size                                  */
#define SR_ERR_HIP           2   /* HIP runtime error (message in sr_last_error())            */
#define SR_ERR_STATE         3   /* operation not valid in the current state (e.g. level)     */
#define SR_ERR_NOMEM         4
#define SR_ERR_NOT_CONVERGED 5   /* solve hit vcCount without meeting its stop rule           */
#define SR_ERR_IO            6   /* file error (restart files, saras_restart.h)               */

/* which level array (lib/poisson/poisson.h:70-73) */
#define SR_PRESSURE 0            /* pressureData */
#define SR_RESIDUAL 1            /* residualData: the right-hand side of each level */
#define SR_TMP      2            /* tmpDataArray: residual f - L u computed by computeResidual */

#define SR_SEM_REFERENCE 0       /* padded-array mean removal, stop on |r_prev - r| < mgTolerance */
#define SR_SEM_PROJECTED 1       /* RHS projected onto range(L), stop on relative residual       */

#define SR_PREC_FP64  0
#define SR_PREC_MIXED 1          /* FP32 inner V-cycles under an FP64 iterative-refinement loop   */
#define SR_PREC_MIXED_GCR 2      /* FP32 V-cycle as the preconditioner of flexible GCR(1) with the
                                    FP64 residual recomputed from the iterate every iteration      */

typedef struct sr_mg sr_mg;

/* Parameters: the multigrid fields of the reference parser (lib/io/parser.h:65-102) plus the
 * grid metrics of lib/grid/grid.cc the solver copies at construction (poisson.cc:453-496).      */
typedef struct sr_mg_desc {
    int xInd, yInd, zInd;        /* N_d = 2^ind + 1 nodes per direction                         */
    int vcDepth, vcCount;
    int preSmooth, postSmooth;
    int gsSmooth;                /* 1: Gauss-Seidel (red-black order on the device); 0: Jacobi   */
    int resType;                 /* computeError norm: 0 none, 1 max, 2 RMS ratio                */
    int xPer, yPer, zPer;
    int testPoisson;             /* TEST_POISSON: Dirichlet ghosts from wall planes, no mean
                                    removal, no early stop (poisson.cc:101-129, 144-147)         */
    int semantics;               /* SR_SEM_*                                                     */
    int precision;               /* SR_PREC_*                                                    */
    int padMode;                 /* 0: reference residual pads for restriction (zero except the
                                    periodic x/y exchange); 1: pads follow each direction's ghost
                                    rule (mirror for Neumann, wrap for periodic z)                 */
    int coarseProject;           /* 1: project the coarsest RHS onto range(L) (singular only)    */
    int coarseDirect;            /* 1: the coarsest level is solved exactly by a precomputed dense
                                    inverse (bordered with the constant vector when L is singular,
                                    which also projects the RHS) instead of Gauss-Seidel to
                                    mgTolerance; at most 1024 coarsest nodes                        */
    double mgTolerance;          /* coarsest-solve absolute tolerance; stagnation tolerance      */
    double relTolerance;         /* SR_SEM_PROJECTED stop: ||f - L u||_2 / ||f||_2 < relTolerance */
    double dXi, dEt, dZt;        /* transformed-plane spacing, lib/grid/grid.cc:87-89           */
    /* staggered metrics of the finest level, host pointers, padded length N+2 (index -1..N):
       xix2Staggr, xixxStaggr, ety2Staggr, etyyStaggr, ztz2Staggr, ztzzStaggr (grid.h)         */
    const double *xix2, *xixx, *ety2, *etyy, *ztz2, *ztzz;
    /* testPoisson only: Dirichlet wall planes, host pointers or NULL, padded 2D over the other two
       directions in (a,b) order: x lo, x hi, y lo, y hi, z lo, z hi (poisson_d3.cc:513-572)      */
    const double *wall[6];
    /* smoothing counts of the FP32 cycle of the mixed precisions (the preconditioner: its counts need
       not be the reference's); 0 or less: preSmooth / postSmooth                                 */
    int innerPreSmooth, innerPostSmooth;
} sr_mg_desc;

typedef struct sr_mg_stats {
    int    cycles;               /* V-cycles (or refinement steps) run                           */
    int    converged;            /* stop rule met before vcCount                                 */
    double residual;             /* last computeError value                                      */
    double solve_ms;             /* wall time of the call, host clock                            */
    double sweep0_ms;            /* mean device time of one finest-level red-black half sweep over
                                    the solve, from HIP events recorded on the solver's stream
                                    around the finest pre-smoothing sweeps of every cycle        */
    double residual0;            /* stop-rule residual of the initial guess, evaluated by the mixed
                                    precisions before their first cycle; -1 for the FP64 path,
                                    whose loop (as the reference's) does not evaluate it         */
} sr_mg_stats;

/* multigrid_d3::multigrid_d3 (lib/poisson/poisson_d3.cc:64-87): sizes, coefficients, metrics,
 * level arrays.  Allocates device memory and a stream.                                           */
int sr_mg_create(const sr_mg_desc *desc, sr_mg **out);
/* poisson::~poisson (lib/poisson/poisson.cc:645-652) */
int sr_mg_destroy(sr_mg *mg);

/* poisson::mgSolve(plainsf &inFn, const plainsf &rhs) (lib/poisson/poisson.cc:100-151).
 * p: in = initial guess (core used), out = full padded solution; rhs: padded, core used.
 * on_device = 0: host pointers; 1: device pointers (HBM-resident, no PCIe transfer).
 * Synchronous with respect to p at return.                                                      */
int sr_mg_solve(sr_mg *mg, double *p, const double *rhs, int on_device, sr_mg_stats *stats);

/* Relative residual of a guess under the solve's own stop rule: computeError
 * (poisson_d3.cc:391-459) evaluated for an arbitrary initial guess p and right-hand side rhs
 * (device pointers, padded layout), loaded exactly as sr_mg_solve loads them.  With r != NULL,
 * the residual field f - L p is written there (padded layout).  Used to judge an initial guess
 * before a solve; the reference evaluates computeError only inside mgSolve.                     */
int sr_mg_residual(sr_mg *mg, const double *p, const double *rhs, double *r, double *rel);

/* Residual after each cycle of the last solve (the values the reference prints with
 * printResidual, poisson.cc:135-142).  Returns the count written (<= cap) or -1.                */
int sr_mg_history(const sr_mg *mg, double *out, int cap);

/* ---- per-operator entry points: the protected virtuals of multigrid_d3 (poisson.h:193-219),
 *      acting on the current vLevel, for parity tests against the oracle ------------------- */
int sr_mg_set_level(sr_mg *mg, int level);                 /* vLevel                          */
int sr_mg_get_level(const sr_mg *mg, int *level);
int sr_mg_set_zero_bc(sr_mg *mg, int zero_bc);             /* zeroBC (poisson.h:49)           */
/* copy a level array between a host padded array (reference layout) and the device          */
int sr_mg_upload(sr_mg *mg, int which, int level, const double *host);
int sr_mg_download(sr_mg *mg, int which, int level, double *host);
int sr_mg_level_dims(const sr_mg *mg, int level, int *nx, int *ny, int *nz);
int sr_mg_op_smooth(sr_mg *mg, int count);                 /* poisson_d3.cc:114-187          */
int sr_mg_op_compute_residual(sr_mg *mg);                  /* poisson_d3.cc:90-111           */
int sr_mg_op_coarsen(sr_mg *mg);                           /* poisson_d3.cc:250-298          */
int sr_mg_op_prolong(sr_mg *mg);                           /* poisson_d3.cc:301-388          */
int sr_mg_op_solve_coarsest(sr_mg *mg);                    /* poisson_d3.cc:190-247          */
int sr_mg_op_impose_bc(sr_mg *mg);                         /* poisson_d3.cc:575-668          */
int sr_mg_op_update_pads(sr_mg *mg, int which);            /* poisson_d3.cc:671-686          */
int sr_mg_op_compute_error(sr_mg *mg, int norm, double *out); /* poisson_d3.cc:391-459       */
int sr_mg_op_vcycle(sr_mg *mg);                            /* poisson.cc:219-280             */

/* ---- measurement ---------------------------------------------------------------------------- */
/* Average device time (ms, HIP events on the solver's stream) of `reps` back-to-back launches of
 * one kernel at one level: kernel 0 = red-black half sweep, 1 = residual, 2 = restriction,
 * 3 = prolongation-correction, 4 = ghost fill, 5 = level-0 residual norm, 6 = the FP32 half
 * sweep of the mixed-precision hierarchy, 7 / 8 = the single-workgroup coarse end of the FP64 /
 * FP32 cycle (level argument ignored).  Device state is modified (scratch use only).           */
int sr_mg_time_kernel(sr_mg *mg, int kernel, int level, int reps, double *avg_ms);
/* Algorithmic bytes one launch of that kernel moves (the roofline numerator).                   */
int sr_mg_kernel_bytes(const sr_mg *mg, int kernel, int level, double *bytes);

/* Plan facts of a handle: what 0 = first level of the one-workgroup coarse cycle (FP64 hierarchy),
 * 1 = same for the FP32 hierarchy (a value above vcDepth means per-level launches throughout),
 * 2 / 3 = LDS bytes of those kernels, 4 = the device's LDS limit per workgroup, 5 = iterations
 * of the last coarsest-level solve (the count solve() reaches, poisson_d3.cc:190-247).            */
int sr_mg_info(const sr_mg *mg, int what, long *value);

/* ---- grid ------------------------------------------------------------------------------------ */
/* One direction of lib/grid/grid.cc: faces = 0 gives the n node ("staggered") coordinates and
 * metrics (padded length n+2, index -1..n); faces = 1 the n-1 face ("collocated") ones (padded
 * length n+1).  grid_type 1 = uniform (createUniformGrid, grid.cc:370-422), 2 = tanh stretching
 * with parameter beta (createTanHypGrid, grid.cc:442-519; pads keep the uniform values).          */
int sr_grid_axis(int n, double len, int grid_type, double beta, int faces, double *x, double *xi_x,
                 double *xixx, double *xix2);

/* ---- device memory ----------------------------------------------------------------------------- */
/* HBM-resident fields for on_device solves: the device counterpart of the Blitz allocation every
 * reference field makes (lib/field/field.cc:75-86).  kind: 1 host->device, 2 device->host,
 * 3 device->device.  Copies complete before return.                                               */
int sr_device_alloc(unsigned long bytes, void **ptr);
int sr_device_free(void *ptr);
int sr_memcpy(void *dst, const void *src, unsigned long bytes, int kind);

/* Last error message of the calling thread (never NULL). */
const char *sr_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
