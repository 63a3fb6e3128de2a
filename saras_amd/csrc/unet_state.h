// State of a learned-guess handle (include/saras_guess.h), shared by the network entry points
// (unet_api.hip) and the solver-facing ones (guess_solver.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/saras_guess.h"
#include "unet_kernels.h"

namespace sr {

struct LayerSpec {
    int cin, cout, level;
};
// e0a e0b | e1a e1b | e2a e2b | b_a b_b | d2 | d1 | d0a out
inline const LayerSpec kLayers[SR_GUESS_LAYERS] = {{8, 16, 0},  {16, 16, 0}, {16, 32, 1}, {32, 32, 1},
                                            {32, 64, 2}, {64, 64, 2}, {64, 64, 3}, {64, 64, 3},
                                            {64, 32, 2}, {32, 16, 1}, {16, 16, 0}, {16, 16, 0}};
struct TensorSpec {
    int level, C;
};
// x0 a0 s0 x1 a1 s1 x2 a2 s2 x3 a3 bb u2 g2 u1 g1 u0 g0
inline const TensorSpec kTensors[18] = {{0, 8},  {0, 16}, {0, 16}, {1, 16}, {1, 32}, {1, 32}, {2, 32}, {2, 64}, {2, 64},
                                 {3, 64}, {3, 64}, {3, 64}, {2, 64}, {2, 32}, {1, 32}, {1, 16}, {0, 16}, {0, 16}};
enum { X0, A0, S0, X1, A1, S1, X2, A2, S2, X3, A3, BB, U2, G2, U1, G1, U0, G0 };

}  // namespace sr

// workgroups of the linear predictor's streaming kernels over the solver grid (dots, combination):
// 16 waves per CU keep enough FP64 loads in flight
constexpr int kLinBlocks = 4096;

struct sr_guess {
    sr_guess_desc d{};
    sr::NetGrid lv[4]{};
    hipStream_t s = nullptr;
    float *W[SR_GUESS_LAYERS] = {}, *b[SR_GUESS_LAYERS] = {};
    float *dW[SR_GUESS_LAYERS] = {}, *db[SR_GUESS_LAYERS] = {};
    float *mW[SR_GUESS_LAYERS] = {}, *vW[SR_GUESS_LAYERS] = {}, *mb[SR_GUESS_LAYERS] = {}, *vb[SR_GUESS_LAYERS] = {};
    sr::bf16_t *pf[SR_GUESS_LAYERS] = {}, *pb[SR_GUESS_LAYERS] = {};
    sr::bf16_t *T[18] = {};
    float *y = nullptr, *dyd = nullptr;
    float *DS0 = nullptr, *DS1 = nullptr, *DS2 = nullptr, *tF = nullptr;
    sr::bf16_t *tM = nullptr, *dy16 = nullptr;
    float *part = nullptr, *bpart = nullptr, *stage = nullptr;
    int adam_t = 0;
    std::vector<void *> owned;
    // solver-facing part: the solver grid, its ratio to the network grid, fine-grid buffers
    int fine[3] = {0, 0, 0};
    int ratio = 1;
    long fine_pad = 0;
    double *extrap = nullptr, *rfield = nullptr;
    float *stage4 = nullptr, *target = nullptr;
    double *rpart = nullptr, *dsc = nullptr, *hsc = nullptr;
    double kappa = 0.0;
    int trained = 0;
    int policy = 0;   // 0: best guess by residual; 1: learned always; 2: adaptive network use
    bool have_cache = false;
    const double *train_base = nullptr;   // base guess of the last predict (training target origin)
    int net_losses = 0, net_skip = 0;     // adaptive policy state
    int net_backoff = 8;
    // linear predictor: a ring of the last hcap converged pairs (projected right-hand side f_i,
    // solution p_i, L p_i = f_i to the solve's tolerance) and their Gram matrix (f_i, f_j)
    int hcap = 0, hn = 0, hnext = 0;
    double *hf[SR_GUESS_MAX_HISTORY] = {}, *hp[SR_GUESS_MAX_HISTORY] = {};
    double G[SR_GUESS_MAX_HISTORY][SR_GUESS_MAX_HISTORY] = {};
    double *fcur = nullptr, *xlin = nullptr;   // projected rhs of the current step, predictor output
    double hd[SR_GUESS_MAX_HISTORY] = {};      // (f_i, fcur)
    double hc[SR_GUESS_MAX_HISTORY] = {};      // least-squares coefficients of the current step
    double yy = 0.0, fit_r2 = 0.0;             // ||fcur||^2 and the fit's squared residual
    bool have_fit = false;                     // fcur / hd / hc belong to the current step
    double *lpart = nullptr, *lsum = nullptr, *lsum_h = nullptr;   // lpart: kLinBlocks x 16 partials
};

namespace sr {
int guess_forward_device(sr_guess *g);
int guess_backward_device(sr_guess *g);
int guess_adam_device(sr_guess *g);
}  // namespace sr

