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
} sr_guess_desc;

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
    double rel_warm;              /* relative residual (sr_mg_residual) of the warm start p1     */
    double rel_extrap;            /* ... of the extrapolation 2 p1 - p2                           */
    double rel_learned;           /* ... of the extrapolation plus the network's correction      */
    int choice;                   /* guess written to p_out: 0 warm start, 1 extrapolation, 2 learned */
    double predict_ms;
} sr_guess_stats;

/* Initial guess for solving L p = rhs with solver m (whose grid is the handle's `fine` grid):
 * p1, p2 = the two previous solutions (p2 may be NULL), all device pointers in the padded layout.
 * Writes the best of {p1, 2 p1 - p2, 2 p1 - p2 + learned correction} by relative residual into
 * p_out.  Keeps the network's forward pass for a following sr_guess_train.                      */
int sr_guess_predict(sr_guess *g, sr_mg *m, const double *p1, const double *p2, const double *rhs,
                     double *p_out, sr_guess_stats *stats);
/* One training step on the converged solution p_star of the solve the last predict was for:
 * target = p_star - (2 p1 - p2) restricted to the network grid; MSE loss, backward pass, Adam.
 * *rel_loss = ||y - target||^2 / ||target||^2 (normalised units) before the step.               */
int sr_guess_train(sr_guess *g, const double *p_star, double *rel_loss);
/* p_out = 2 p1 - p2 (p1 when p2 is NULL): the extrapolated guess alone, for comparisons */
int sr_guess_extrapolate(sr_guess *g, const double *p1, const double *p2, double *p_out);
/* policy 0 (default): sr_guess_predict returns the best of the three guesses by relative
 * residual; 1: always the learned one (ablation and diagnostics)                               */
int sr_guess_set_policy(sr_guess *g, int policy);
/* the correction scale kappa and the number of training steps taken */
int sr_guess_kappa(const sr_guess *g, double *kappa, int *trained);

#ifdef __cplusplus
}
#endif
#endif
