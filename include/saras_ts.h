/*
 * saras_ts.h -- C ABI of the MI355X hydro time step: the field stencils, the implicit Jacobi
 * viscous solve and the pressure projection that src/solvers/hydro/hydro_d3.cc runs every step,
 * with the multigrid pressure solve of saras_mg.h inside.
 *
 * Fields use the reference's padded layout (C order, k fastest, index -1 first):
 *   Vx: (Nx-1+2) x (Ny+2) x (Nz+2)   (x faces, y/z nodes; lib/field/vfield.cc:64-66)
 *   Vy: (Nx+2) x (Ny-1+2) x (Nz+2),  Vz: (Nx+2) x (Ny+2) x (Nz-1+2)
 *   P, Pp, mgRHS, divergence: (Nx+2) x (Ny+2) x (Nz+2)
 * All fields live in HBM for the life of the handle.  Status codes as in saras_mg.h.
 */
#ifndef SARAS_TS_H
#define SARAS_TS_H

#include "saras_guess.h"
#include "saras_mg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* which field */
#define SR_TS_VX 0
#define SR_TS_VY 1
#define SR_TS_VZ 2
#define SR_TS_P 3
#define SR_TS_PP 4          /* pressure correction (the multigrid solution, warm start) */
#define SR_TS_MGRHS 5
#define SR_TS_HX 6          /* nseRHS components */
#define SR_TS_HY 7
#define SR_TS_HZ 8
#define SR_TS_DIV 12
#define SR_TS_T 13          /* temperature (scalar runs) */
#define SR_TS_TRHS 14       /* its right-hand side tmpRHS */
#define SR_TS_PP1 15        /* the previous two pressure corrections (kept with a guess attached) */
#define SR_TS_PP2 16

typedef struct sr_ts sr_ts;

/* The parser fields the time step reads (lib/io/parser.h:65-102). */
typedef struct sr_ts_desc {
    int xInd, yInd, zInd;
    double Lx, Ly, Lz;
    int xGrid, yGrid, zGrid;        /* 1 uniform, 2 tanh */
    double betaX, betaY, betaZ;
    int xPer, yPer, zPer;
    int probType;                   /* velocity walls of hydro::initVBC (hydro.cc:287-371) */
    double Re, dt, cnTolerance;
    int vcDepth, vcCount, preSmooth, postSmooth, resType, gsSmooth;
    double mgTolerance;
    int scalar;                     /* 1: scalar_d3 (temperature, buoyancy forcing, forceType 3)  */
    int rbcType;                    /* scalar.cc:69-86 non-dimensionalisation                      */
    double Ra, Pr;
    int semantics;                  /* pressure solve stop rule: SR_SEM_REFERENCE / _PROJECTED     */
    double relTolerance;
    int precision;                  /* pressure solve: SR_PREC_FP64, _MIXED or _MIXED_GCR          */
    int coarseDirect;               /* pressure solve: sr_mg_desc.coarseDirect                     */
    int innerPreSmooth, innerPostSmooth; /* pressure solve: sr_mg_desc.innerPreSmooth / _PostSmooth */
    int padMode;                    /* pressure solve: sr_mg_desc.padMode                          */
} sr_ts_desc;

typedef struct sr_ts_stats {
    int jacobi_iters[4];            /* sweeps of solveVx / solveVy / solveVz / solveT              */
    int mg_cycles;                  /* V-cycles of the pressure solve                              */
    double step_ms;                 /* wall time of the call                                       */
    double solve_ms;                /* wall time of the pressure solve (guess excluded)            */
    double guess_ms;                /* wall time of the guess: sr_guess_predict before the solve,  */
                                    /* sr_guess_train, sr_guess_observe and the copy of the solution */
                                    /* kept as its next input after it (0 without a guess)         */
    int guess_choice;               /* sr_guess_stats.choice, -1 without a guess                   */
    double rel_guess;               /* relative residual of the guess handed to the solve          */
    double train_loss;              /* sr_guess_train's relative loss, -1 when not training        */
    double sweep0_ms;               /* sr_mg_stats.sweep0_ms of the pressure solve                 */
    int mg_converged;               /* 1: the pressure solve met its stop rule; 0: it ran vcCount  */
                                    /* cycles without meeting it and the step went on with the last */
                                    /* iterate, as mgSolve does (poisson.cc:139-149)               */
    double mg_residual;             /* sr_mg_stats.residual of the pressure solve                  */
    double guess_rel_linear;        /* sr_guess_stats.rel_linear: the linear predictor's relative  */
                                    /* residual (exact when the network ran, else its fit's estimate), */
                                    /* -1 without one                                              */
    double guess_rel_learned;       /* sr_guess_stats.rel_learned, -1 when the network did not run */
    double guess_predict_ms;        /* sr_guess_stats.predict_ms: the part of guess_ms before the solve */
} sr_ts_stats;

/* hydro_d3::hydro_d3 (src/solvers/hydro/hydro_d3.cc:63-119) without the initial condition */
int sr_ts_create(const sr_ts_desc *desc, sr_ts **out);
int sr_ts_destroy(sr_ts *ts);
/* P = 1 and the Taylor-Green velocity (lib/initial/taylorGreen.cc:54-100), then the velocity BCs */
int sr_ts_init_tgv(sr_ts *ts);
/* host <-> device copies of a whole padded field */
int sr_ts_upload(sr_ts *ts, int which, const double *host);
int sr_ts_download(sr_ts *ts, int which, double *host);
int sr_ts_field_dims(const sr_ts *ts, int which, int *nx, int *ny, int *nz);   /* core sizes */
/* the device array of a field, in the padded layout (for probes, sr_probe_gather, and for the field
   operators of saras_field.h on the handle's fields); valid until sr_ts_destroy */
int sr_ts_field_ptr(sr_ts *ts, int which, double **dev);

/* hydro_d3::timeAdvance (hydro_d3.cc:254-368), zero forcing; with desc.scalar the step of
 * scalar_d3::timeAdvance (scalar_d3.cc:243-370) with buoyancy                                   */
int sr_ts_time_advance(sr_ts *ts, sr_ts_stats *stats);

/* ---- pieces of the step, for parity tests ------------------------------------------------- */
/* derivative::calcDerivative1_{x,y,z} / calcDerivative2{xx,yy,zz} (lib/field/derivative.cc:121-243)
 * of field `which` (a velocity component or P) along `dim`; out is a host array of that field's
 * padded shape, written over the core range along dim and every index of the other two.           */
int sr_ts_deriv(sr_ts *ts, int which, int order, int dim, double *out);
/* nseRHS = V + dt (computeDiff / Re - computeNLin - grad P), hydro_d3.cc:259-303 (core values) */
int sr_ts_build_rhs(sr_ts *ts);
/* solveVx / solveVy / solveVz (hydro_d3.cc:370-533); iterations returned in *iters              */
int sr_ts_solve_v(sr_ts *ts, int component, int *iters);
/* hydro::imposeU/V/WBCs (hydro.cc:386-461) */
int sr_ts_impose_vbc(sr_ts *ts, int component);
/* vfield::divergence into SR_TS_DIV (vfield.cc:299-320) */
int sr_ts_divergence(sr_ts *ts);
/* tseries energy and maximum divergence (lib/io/tseries.cc:137-178) */
int sr_ts_energy(sr_ts *ts, double *energy);
int sr_ts_max_div(sr_ts *ts, double *maxdiv);

/* With a guess handle attached (its solver grid = this run's node grid), each pressure solve
 * starts from sr_guess_predict's guess built from the previous two corrections and the new
 * right-hand side, and with train != 0 the network takes one training step on every converged
 * correction.  g = NULL detaches (the reference's warm start from the previous correction).
 * No counterpart in the reference.                                                               */
int sr_ts_attach_guess(sr_ts *ts, sr_guess *g, int train);
/* The handle's pressure solver (owned by the time step; valid until sr_ts_destroy), e.g. for
 * sr_mg_history of the last solve or sr_mg_info.                                                */
int sr_ts_pressure_solver(sr_ts *ts, sr_mg **mg);

/* Measurement (no reference counterpart): the algorithmic HBM bytes of one launch of a step kernel
   and its average duration over `reps` launches on the handle's fields (HIP events on the handle's
   stream; the kernels overwrite what they write, so a handle timed this way is for measurement).
   kernel 0: the three nseRHS kernels (40 B per face node: V_c, the two other components, P read,
   H_c written); 1: one Jacobi sweep of Vx (24 B per padded entry); 2: its residual maximum (16 B per
   bulk node); 3: divergence (32 B per node); 4: the three projection updates (24 B per face node);
   5: the energy reduction (24 B per node); 6: one Jacobi sweep of Vx together with the residual
   maximum of its input, the form the step runs (24 B per padded entry). */
int sr_ts_kernel_bytes(const sr_ts *ts, int kernel, double *bytes);
int sr_ts_time_kernel(sr_ts *ts, int kernel, int reps, double *avg_ms);

/* ---- scalar_d3 (src/solvers/scalar/scalar_d3.cc) ------------------------------------------- */
/* the constructor's state (scalar_d3.cc:85-107: P = 1, T = 0, V = 0) with the temperature
 * perturbed on the core by amp (2u - 1), u uniform from SplitMix64(seed) in C order; then the
 * velocity and temperature BCs                                                                   */
int sr_ts_init_rbc(sr_ts *ts, unsigned long long seed, double amp);
/* scalar::imposeTBCs (scalar.cc:240-255) with the walls of initTBC for probType 5/6              */
int sr_ts_impose_tbc(sr_ts *ts);
/* tmpRHS = T + dt (kappa computeDiff(T) - computeNLin(V, T)), synced (scalar_d3.cc:349-360)     */
int sr_ts_build_trhs(sr_ts *ts);
/* scalar_d3::solveT (scalar_d3.cc:533-594)                                                       */
int sr_ts_solve_t(sr_ts *ts, int *iters);
/* tseries Nusselt number 1 + <w T>/kappa (tseries.cc:195-246)                                    */
int sr_ts_nusselt(sr_ts *ts, double *nu);
/* the run's kinematic viscosity and thermal diffusivity: hydro nu = 1/Re (hydro.cc), scalar runs by
 * rbcType (scalar.cc:69-86); tseries reports sqrt(E)/nu for scalar runs (tseries.cc:240). kappa is
 * 0 for a hydro run.                                                                             */
int sr_ts_coefficients(sr_ts *ts, double *nu, double *kappa);

#ifdef __cplusplus
}
#endif
#endif
