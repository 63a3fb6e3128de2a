/*
 * saras_guess.h -- C ABI of the learned initial guess for the pressure solve.
 *
 * No counterpart in the reference: SARAS starts each mgSolve from the previous step's pressure
 * correction (src/solvers/hydro/hydro_d3.cc:343, Pp passed in/out).  These entry points add a
 * small 3D U-Net, run and trained on the GPU by its own bf16 matrix-core kernels, whose output
 * corrects an extrapolated initial guess; the multigrid solve still iterates to its tolerance,
 * and the guess is used only when it lowers the initial residual (saras_mg.h, sr_mg_solve).
 *
 * Network (layer order of sr_guess_set_layer): 4 levels of node-centred grids (n, (n-1)/2+1, ...),
 * channels 16/32/64/64, two 3x3x3 convolutions + ReLU per encoder level, full-weighting
 * restriction down, trilinear prolongation up with additive skips, one convolution per decoder
 * level, a linear 1-channel output.  Weights W[27][cout][cin] (tap = (dx+1)*9+(dy+1)*3+(dz+1)),
 * biases b[cout]; fp32 master copies, bf16 on the matrix cores, fp32 accumulation.
 * Tensors are channels-last: value (x, y, z, c) at ((x*ny + y)*nz + z)*C + c.
 */
#ifndef SARAS_GUESS_H
#define SARAS_GUESS_H

#include "saras_mg.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SR_GUESS_LAYERS 12
#define SR_GUESS_IN_CH 8          /* stored input channels (3 used) */

typedef struct sr_guess sr_guess;

typedef struct sr_guess_desc {
    int n[3];                     /* level-0 nodes per direction, each 2^k + 1 with k >= 3     */
    int fine[3];                  /* solver grid nodes: n (or 0) or 2n - 1 in every direction  */
    unsigned long long seed;      /* initial weights (He normal, zero biases)                   */
    float lr, beta1, beta2, eps;  /* Adam                                                        */
    int history;                  /* 0..SR_GUESS_MAX_HISTORY: converged (rhs, solution) pairs kept
                                     for the linear least-squares predictor (0: off)              */
} sr_guess_desc;

#define SR_GUESS_MAX_HISTORY 8

int sr_guess_create(const sr_guess_desc *desc, sr_guess **out);
int sr_guess_destroy(sr_guess *g);
/* layer l: its input/output channel counts and level */
int sr_guess_layer_info(int layer, int *cin, int *cout, int *level);
/* fp32 master weights and biases of one layer (host arrays W[27*cout*cin], b[cout]) */
int sr_guess_set_layer(sr_guess *g, int layer, const float *W, const float *b);
int sr_guess_get_layer(sr_guess *g, int layer, float *W, float *b);
/* gradients of the last sr_guess_backward */
int sr_guess_get_grad(sr_guess *g, int layer, float *dW, float *db);

/* y = net(x0): x0 host [n^3 * SR_GUESS_IN_CH] (rounded to bf16), y host [n^3] fp32 */
int sr_guess_forward(sr_guess *g, const float *x0, float *y);
/* gradients of sum(dy * y) for the last forward, dy host [n^3] */
int sr_guess_backward(sr_guess *g, const float *dy);
/* one Adam step on the last gradients */
int sr_guess_adam(sr_guess *g);
/* stored tensor `which` of the last forward as fp32 (0 x0, 1 a0, 2 s0, 3 x1, 4 a1, 5 s1, 6 x2,
 * 7 a2, 8 s2, 9 x3, 10 a3, 11 bb, 12 u2, 13 g2, 14 u1, 15 g1, 16 u0, 17 g0); *count = elements */
int sr_guess_tensor(sr_guess *g, int which, float *host, long *count);

/* ---- use inside a sequence of pressure solves --------------------------------------------- */
typedef struct sr_guess_stats {
    double rel_warm;              /* relative residual (sr_mg_residual) of the warm start p1, or -1 */
    double rel_extrap;            /* ... of the extrapolation 2 p1 - p2, or -1                    */
    double rel_learned;           /* ... of the base guess plus the network's correction, or -1  */
    int choice;                   /* guess written to p_out: 0 warm start, 1 extrapolation,
                                     2 learned (base + network correction), 3 linear predictor   */
    double predict_ms;
    double rel_linear;            /* relative residual of the linear predictor (from its
                                     least-squares fit; exact when the network ran), or -1      */
    int net_used;                 /* 1 when the network ran for this guess                       */
} sr_guess_stats;

/* Initial guess for solving L p = rhs with solver m (whose grid is the handle's `fine` grid):
 * p1, p2 = the two previous solutions (p2 may be NULL), all device pointers in the padded layout.
 * Base guess: with a history (desc.history > 0, projected semantics, at least one pair observed)
 * the linear predictor -- the least-squares combination of the observed solutions whose right-hand
 * sides best fit rhs (Gram-Schmidt basis of the projected right-hand sides; its residual is known
 * from the fit) -- which contains p1 and 2 p1 - p2 in its span; otherwise 2 p1 - p2 (or p1 when
 * its residual is lower).  The network's correction is added to the base and kept only when its
 * relative residual (sr_mg_residual) is lower.  Keeps the forward pass for sr_guess_train.       */
int sr_guess_predict(sr_guess *g, sr_mg *m, const double *p1, const double *p2, const double *rhs,
                     double *p_out, sr_guess_stats *stats);
/* One training step on the converged solution p_star of the solve the last predict was for:
 * target = p_star - (base guess) restricted to the network grid; MSE loss, backward pass, Adam.
 * *rel_loss = ||y - target||^2 / ||target||^2 (normalised units) before the step; -1 when the
 * last predict did not run the network (nothing to learn from).                                 */
int sr_guess_train(sr_guess *g, const double *p_star, double *rel_loss);
/* Add the converged pair (rhs, p_star) of the solve just finished to the linear predictor's
 * history (call after every solve, trained or not; a no-op without a history).  When the history
 * is full it keeps the newest half of its basis and adds this pair.                            */
int sr_guess_observe(sr_guess *g, sr_mg *m, const double *rhs, const double *p_star);
/* p_out = 2 p1 - p2 (p1 when p2 is NULL): the extrapolated guess alone, for comparisons */
int sr_guess_extrapolate(sr_guess *g, const double *p1, const double *p2, double *p_out);
/* policy 0 (default): sr_guess_predict runs the network every time and returns the best guess by
 * relative residual; 1: always the learned one (ablation and diagnostics); 2: adaptive -- after
 * three predictions in a row where the network's guess lost, the network is skipped for the next
 * eight, and after each further losing probe for twice as many (up to 256; a win resets it); the
 * record of policy 0 carries over, so a network that lost its last three comparisons in training
 * starts skipped (for a frozen network; training needs policy 0); 3: never the network (the base
 * guess alone: linear predictor or extrapolation; ablation)                                      */
int sr_guess_set_policy(sr_guess *g, int policy);
/* the correction scale kappa and the number of training steps taken */
int sr_guess_kappa(const sr_guess *g, double *kappa, int *trained);
/* Measurement: the average duration of one layer's kernel over `reps` launches on the handle's
   current tensors (HIP events on the handle's stream; the work tensors it writes are overwritten).
   which 0: the convolution of the forward pass; 1: its adjoint (the data gradient); 2: the weight
   gradient with the bias gradient.  flops = 2 * 27 * cin * cout * voxels of the layer's grid, the
   multiply-adds of one 3x3x3 convolution (the same count for the three). */
int sr_guess_time_layer(sr_guess *g, int layer, int which, int reps, double *avg_ms, double *flops);

#ifdef __cplusplus
}
#endif
#endif
