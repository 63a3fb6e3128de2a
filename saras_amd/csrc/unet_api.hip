// C ABI of the learned initial guess (include/saras_guess.h): the U-Net's state in HBM, its
// forward and backward passes over the kernels of unet_kernels.hip, and Adam.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/saras_guess.h"
#include "mg_state.h"
#include "unet_kernels.h"
#include "unet_state.h"

using namespace sr;

#define HIPCHK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t _e = (expr);                                                                       \
        if (_e != hipSuccess)                                                                         \
            return set_err(SR_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));            \
    } while (0)

#define RETCHK(expr)                                                                                  \
    do {                                                                                              \
        int _r = (expr);                                                                              \
        if (_r != SR_OK) return _r;                                                                   \
    } while (0)

static long tensor_elems(const sr_guess *g, int t) { return g->lv[kTensors[t].level].nvox() * kTensors[t].C; }
static long wsize(int l) { return 27L * kLayers[l].cin * kLayers[l].cout; }

template <typename T>
static int dalloc(sr_guess *g, T **p, long n) {
    HIPCHK(hipMalloc((void **)p, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    g->owned.push_back((void *)*p);
    HIPCHK(hipMemsetAsync(*p, 0, sizeof(T) * (size_t)(n > 0 ? n : 1), g->s));
    return SR_OK;
}

static void repack(sr_guess *g, int l) {
    launch_pack(g->W[l], kLayers[l].cin, kLayers[l].cout, 0, g->pf[l], g->s);
    if (l > 0) launch_pack(g->W[l], kLayers[l].cin, kLayers[l].cout, 1, g->pb[l], g->s);
}

// SplitMix64 -> uniform -> Box-Muller
static double rng_uniform(uint64_t &st) {
    st += 0x9E3779B97F4A7C15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

static int create_impl(const sr_guess_desc *desc, sr_guess *g) {
    g->d = *desc;
    if (desc->history < 0 || desc->history > SR_GUESS_MAX_HISTORY)
        return set_err(SR_ERR_ARG, "history must be 0..SR_GUESS_MAX_HISTORY");
    for (int k = 0; k < 3; k++) {
        int n = desc->n[k];
        int m = n - 1;
        if (n < 9 || (m & (m - 1)) != 0) return set_err(SR_ERR_ARG, "network grid must have 2^k + 1 nodes, k >= 3");
    }
    // solver grid: equal to the network grid, or twice as fine (2n - 1 nodes) in every direction
    for (int k = 0; k < 3; k++) g->fine[k] = desc->fine[k] > 0 ? desc->fine[k] : desc->n[k];
    g->ratio = g->fine[0] == desc->n[0] ? 1 : 2;
    for (int k = 0; k < 3; k++)
        if (g->fine[k] != (g->ratio == 1 ? desc->n[k] : 2 * desc->n[k] - 1))
            return set_err(SR_ERR_ARG, "solver grid must equal the network grid or have 2n - 1 nodes in every direction");
    int n[3] = {desc->n[0], desc->n[1], desc->n[2]};
    for (int l = 0; l < 4; l++) {
        g->lv[l] = net_grid(n[0], n[1], n[2]);
        for (int k = 0; k < 3; k++) n[k] = (n[k] - 1) / 2 + 1;
    }
    if (g->d.lr <= 0.f) g->d.lr = 1e-3f;
    if (g->d.beta1 <= 0.f) g->d.beta1 = 0.9f;
    if (g->d.beta2 <= 0.f) g->d.beta2 = 0.999f;
    if (g->d.eps <= 0.f) g->d.eps = 1e-8f;
    HIPCHK(hipStreamCreateWithFlags(&g->s, hipStreamNonBlocking));
    long partmax = 0, bpartmax = 0;
    for (int l = 0; l < SR_GUESS_LAYERS; l++) {
        const LayerSpec &L = kLayers[l];
        RETCHK(dalloc(g, &g->W[l], wsize(l)));
        RETCHK(dalloc(g, &g->b[l], L.cout));
        RETCHK(dalloc(g, &g->dW[l], wsize(l)));
        RETCHK(dalloc(g, &g->db[l], L.cout));
        RETCHK(dalloc(g, &g->mW[l], wsize(l)));
        RETCHK(dalloc(g, &g->vW[l], wsize(l)));
        RETCHK(dalloc(g, &g->mb[l], L.cout));
        RETCHK(dalloc(g, &g->vb[l], L.cout));
        RETCHK(dalloc(g, &g->pf[l], conv_pack_elems(L.cin, L.cout)));
        RETCHK(dalloc(g, &g->pb[l], conv_pack_elems(L.cout, L.cin)));
        partmax = std::max(partmax, wgrad_part_elems(L.cin, L.cout, g->lv[L.level]));
        bpartmax = std::max(bpartmax, (long)wgrad_slots(L.cin, L.cout, g->lv[L.level]) * L.cout);
    }
    for (int t = 0; t < 18; t++) RETCHK(dalloc(g, &g->T[t], tensor_elems(g, t)));
    const long n0 = g->lv[0].nvox(), n1 = g->lv[1].nvox(), n2 = g->lv[2].nvox();
    RETCHK(dalloc(g, &g->y, n0));
    RETCHK(dalloc(g, &g->dyd, n0));
    RETCHK(dalloc(g, &g->DS0, n0 * 16));
    RETCHK(dalloc(g, &g->DS1, n1 * 32));
    RETCHK(dalloc(g, &g->DS2, n2 * 64));
    long tmax = std::max(n0 * 16, std::max(n1 * 32, n2 * 64));
    RETCHK(dalloc(g, &g->tF, tmax));
    RETCHK(dalloc(g, &g->tM, tmax));
    RETCHK(dalloc(g, &g->dy16, n0 * 16));
    RETCHK(dalloc(g, &g->part, partmax));
    RETCHK(dalloc(g, &g->bpart, bpartmax));
    RETCHK(dalloc(g, &g->stage, n0 * SR_GUESS_IN_CH));
    g->fine_pad = (long)(g->fine[0] + 2) * (g->fine[1] + 2) * (g->fine[2] + 2);
    RETCHK(dalloc(g, &g->extrap, g->fine_pad));
    RETCHK(dalloc(g, &g->rfield, g->fine_pad));
    RETCHK(dalloc(g, &g->stage4, n0 * 4));
    RETCHK(dalloc(g, &g->target, n0));
    RETCHK(dalloc(g, &g->rpart, 4096L * 4));
    RETCHK(dalloc(g, &g->dsc, 16));
    HIPCHK(hipHostMalloc((void **)&g->hsc, sizeof(double) * 16, hipHostMallocDefault));
    g->hcap = desc->history;
    if (g->hcap > 0) {
        for (int i = 0; i < g->hcap; i++) {
            RETCHK(dalloc(g, &g->hf[i], g->fine_pad));
            RETCHK(dalloc(g, &g->hp[i], g->fine_pad));
        }
        RETCHK(dalloc(g, &g->fcur, g->fine_pad));
        RETCHK(dalloc(g, &g->xlin, g->fine_pad));
        RETCHK(dalloc(g, &g->lpart, (long)kLinBlocks * 16));
        RETCHK(dalloc(g, &g->lsum, 16));
        HIPCHK(hipHostMalloc((void **)&g->lsum_h, sizeof(double) * 16, hipHostMallocDefault));
    }
    // He-normal initial weights (zero biases); the unused input channels and output channels
    // start at zero, the output layer small
    uint64_t st = desc->seed;
    for (int l = 0; l < SR_GUESS_LAYERS; l++) {
        const LayerSpec &L = kLayers[l];
        std::vector<float> w((size_t)wsize(l));
        double sd = std::sqrt(2.0 / (27.0 * L.cin));
        for (size_t e = 0; e < w.size(); e++) {
            double u1 = rng_uniform(st), u2 = rng_uniform(st);
            double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
            int ci = (int)(e % L.cin), co = (int)((e / L.cin) % L.cout);
            if (l == 0 && ci >= 3) z = 0.0;
            if (l == SR_GUESS_LAYERS - 1) z = co == 0 ? 0.1 * z : 0.0;
            w[e] = (float)(z * sd);
        }
        HIPCHK(hipMemcpyAsync(g->W[l], w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice, g->s));
        HIPCHK(hipStreamSynchronize(g->s));
        repack(g, l);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g->s));
    return SR_OK;
}

extern "C" int sr_guess_destroy(sr_guess *g) {
    if (!g) return SR_OK;
    if (g->s) (void)hipStreamSynchronize(g->s);
    for (void *p : g->owned) (void)hipFree(p);
    if (g->hsc) (void)hipHostFree(g->hsc);
    if (g->lsum_h) (void)hipHostFree(g->lsum_h);
    if (g->s) (void)hipStreamDestroy(g->s);
    delete g;
    return SR_OK;
}

extern "C" int sr_guess_create(const sr_guess_desc *desc, sr_guess **out) {
    if (!desc || !out) return set_err(SR_ERR_ARG, "null argument");
    *out = nullptr;
    sr_guess *g = new sr_guess();
    int r = create_impl(desc, g);
    if (r != SR_OK) {
        std::string msg = sr_last_error();
        sr_guess_destroy(g);
        return set_err(r, msg);
    }
    *out = g;
    return SR_OK;
}

extern "C" int sr_guess_layer_info(int layer, int *cin, int *cout, int *level) {
    if (layer < 0 || layer >= SR_GUESS_LAYERS) return set_err(SR_ERR_ARG, "layer out of range");
    if (cin) *cin = kLayers[layer].cin;
    if (cout) *cout = kLayers[layer].cout;
    if (level) *level = kLayers[layer].level;
    return SR_OK;
}

extern "C" int sr_guess_set_layer(sr_guess *g, int l, const float *W, const float *b) {
    if (!g || !W || !b || l < 0 || l >= SR_GUESS_LAYERS) return set_err(SR_ERR_ARG, "bad argument");
    HIPCHK(hipMemcpyAsync(g->W[l], W, sizeof(float) * wsize(l), hipMemcpyHostToDevice, g->s));
    HIPCHK(hipMemcpyAsync(g->b[l], b, sizeof(float) * kLayers[l].cout, hipMemcpyHostToDevice, g->s));
    repack(g, l);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g->s));
    return SR_OK;
}

extern "C" int sr_guess_get_layer(sr_guess *g, int l, float *W, float *b) {
    if (!g || !W || !b || l < 0 || l >= SR_GUESS_LAYERS) return set_err(SR_ERR_ARG, "bad argument");
    HIPCHK(hipMemcpyAsync(W, g->W[l], sizeof(float) * wsize(l), hipMemcpyDeviceToHost, g->s));
    HIPCHK(hipMemcpyAsync(b, g->b[l], sizeof(float) * kLayers[l].cout, hipMemcpyDeviceToHost, g->s));
    HIPCHK(hipStreamSynchronize(g->s));
    return SR_OK;
}

extern "C" int sr_guess_get_grad(sr_guess *g, int l, float *dW, float *db) {
    if (!g || !dW || !db || l < 0 || l >= SR_GUESS_LAYERS) return set_err(SR_ERR_ARG, "bad argument");
    HIPCHK(hipMemcpyAsync(dW, g->dW[l], sizeof(float) * wsize(l), hipMemcpyDeviceToHost, g->s));
    HIPCHK(hipMemcpyAsync(db, g->db[l], sizeof(float) * kLayers[l].cout, hipMemcpyDeviceToHost, g->s));
    HIPCHK(hipStreamSynchronize(g->s));
    return SR_OK;
}

// ------------------------------------------------------------------------------------------------
// forward / backward on the device
// ------------------------------------------------------------------------------------------------
static int conv(sr_guess *g, int l, const bf16_t *in, void *out, int mode) {
    const LayerSpec &L = kLayers[l];
    if (launch_conv(L.cin, L.cout, g->lv[L.level], in, g->pf[l], g->b[l], out, mode, g->s) != 0)
        return set_err(SR_ERR_ARG, "no convolution kernel for this channel pair");
    return SR_OK;
}

// adjoint convolution of layer l: dX (fp32, cin channels) from the bf16 gradient dY (cout channels)
static int conv_adj(sr_guess *g, int l, const bf16_t *dY, float *dX) {
    const LayerSpec &L = kLayers[l];
    if (launch_conv(L.cout, L.cin, g->lv[L.level], dY, g->pb[l], nullptr, dX, CONV_F32, g->s) != 0)
        return set_err(SR_ERR_ARG, "no adjoint convolution kernel for this channel pair");
    return SR_OK;
}

static int wgrad(sr_guess *g, int l, const bf16_t *dY, const bf16_t *X) {
    const LayerSpec &L = kLayers[l];
    // the bias gradient comes out of the same pass over dY
    if (launch_wgrad(L.cin, L.cout, g->lv[L.level], dY, X, g->part, g->dW[l], g->bpart, g->db[l], g->s) != 0)
        return set_err(SR_ERR_ARG, "no weight-gradient kernel for this channel pair");
    return SR_OK;
}

int sr::guess_forward_device(sr_guess *g) {
    const NetGrid *v = g->lv;
    RETCHK(conv(g, 0, g->T[X0], g->T[A0], CONV_BF16_RELU));
    RETCHK(conv(g, 1, g->T[A0], g->T[S0], CONV_BF16_RELU));
    launch_restrict_bf16(v[0], v[1], 16, g->T[S0], g->T[X1], g->s);
    RETCHK(conv(g, 2, g->T[X1], g->T[A1], CONV_BF16_RELU));
    RETCHK(conv(g, 3, g->T[A1], g->T[S1], CONV_BF16_RELU));
    launch_restrict_bf16(v[1], v[2], 32, g->T[S1], g->T[X2], g->s);
    RETCHK(conv(g, 4, g->T[X2], g->T[A2], CONV_BF16_RELU));
    RETCHK(conv(g, 5, g->T[A2], g->T[S2], CONV_BF16_RELU));
    launch_restrict_bf16(v[2], v[3], 64, g->T[S2], g->T[X3], g->s);
    RETCHK(conv(g, 6, g->T[X3], g->T[A3], CONV_BF16_RELU));
    RETCHK(conv(g, 7, g->T[A3], g->T[BB], CONV_BF16_RELU));
    launch_prolong_add_bf16(v[2], v[3], 64, g->T[BB], g->T[S2], g->T[U2], g->s);
    RETCHK(conv(g, 8, g->T[U2], g->T[G2], CONV_BF16_RELU));
    launch_prolong_add_bf16(v[1], v[2], 32, g->T[G2], g->T[S1], g->T[U1], g->s);
    RETCHK(conv(g, 9, g->T[U1], g->T[G1], CONV_BF16_RELU));
    launch_prolong_add_bf16(v[0], v[1], 16, g->T[G1], g->T[S0], g->T[U0], g->s);
    RETCHK(conv(g, 10, g->T[U0], g->T[G0], CONV_BF16_RELU));
    RETCHK(conv(g, 11, g->T[G0], g->y, CONV_F32_CH0));
    return SR_OK;
}

// gradients of sum(dyd * y); the mask of each ReLU is the stored post-activation (> 0)
int sr::guess_backward_device(sr_guess *g) {
    const NetGrid *v = g->lv;
    hipStream_t s = g->s;
    const long n0 = v[0].nvox(), n1 = v[1].nvox(), n2 = v[2].nvox(), n3 = v[3].nvox();
    launch_expand_ch0(n0, g->dyd, g->dy16, s);
    RETCHK(wgrad(g, 11, g->dy16, g->T[G0]));
    RETCHK(conv_adj(g, 11, g->dy16, g->tF));                 // d/dg0
    launch_mask(n0 * 16, g->tF, g->T[G0], g->tM, s);
    RETCHK(wgrad(g, 10, g->tM, g->T[U0]));
    RETCHK(conv_adj(g, 10, g->tM, g->DS0));                  // d/du0 = d/ds0 (skip part)
    launch_restrict_f32(v[0], v[1], 16, g->DS0, g->tF, 8.0f, s);   // prolongation adjoint = 8 R
    launch_mask(n1 * 16, g->tF, g->T[G1], g->tM, s);
    RETCHK(wgrad(g, 9, g->tM, g->T[U1]));
    RETCHK(conv_adj(g, 9, g->tM, g->DS1));                   // d/du1
    launch_restrict_f32(v[1], v[2], 32, g->DS1, g->tF, 8.0f, s);
    launch_mask(n2 * 32, g->tF, g->T[G2], g->tM, s);
    RETCHK(wgrad(g, 8, g->tM, g->T[U2]));
    RETCHK(conv_adj(g, 8, g->tM, g->DS2));                   // d/du2
    launch_restrict_f32(v[2], v[3], 64, g->DS2, g->tF, 8.0f, s);
    launch_mask(n3 * 64, g->tF, g->T[BB], g->tM, s);
    RETCHK(wgrad(g, 7, g->tM, g->T[A3]));
    RETCHK(conv_adj(g, 7, g->tM, g->tF));                    // d/da3
    launch_mask(n3 * 64, g->tF, g->T[A3], g->tM, s);
    RETCHK(wgrad(g, 6, g->tM, g->T[X3]));
    RETCHK(conv_adj(g, 6, g->tM, g->tF));                    // d/dx3
    launch_prolong_add_f32(v[2], v[3], 64, g->tF, g->DS2, g->DS2, 0.125f, s);   // += R^T = P / 8
    launch_mask(n2 * 64, g->DS2, g->T[S2], g->tM, s);
    RETCHK(wgrad(g, 5, g->tM, g->T[A2]));
    RETCHK(conv_adj(g, 5, g->tM, g->tF));
    launch_mask(n2 * 64, g->tF, g->T[A2], g->tM, s);
    RETCHK(wgrad(g, 4, g->tM, g->T[X2]));
    RETCHK(conv_adj(g, 4, g->tM, g->tF));                    // d/dx2
    launch_prolong_add_f32(v[1], v[2], 32, g->tF, g->DS1, g->DS1, 0.125f, s);
    launch_mask(n1 * 32, g->DS1, g->T[S1], g->tM, s);
    RETCHK(wgrad(g, 3, g->tM, g->T[A1]));
    RETCHK(conv_adj(g, 3, g->tM, g->tF));
    launch_mask(n1 * 32, g->tF, g->T[A1], g->tM, s);
    RETCHK(wgrad(g, 2, g->tM, g->T[X1]));
    RETCHK(conv_adj(g, 2, g->tM, g->tF));                    // d/dx1
    launch_prolong_add_f32(v[0], v[1], 16, g->tF, g->DS0, g->DS0, 0.125f, s);
    launch_mask(n0 * 16, g->DS0, g->T[S0], g->tM, s);
    RETCHK(wgrad(g, 1, g->tM, g->T[A0]));
    RETCHK(conv_adj(g, 1, g->tM, g->tF));
    launch_mask(n0 * 16, g->tF, g->T[A0], g->tM, s);
    RETCHK(wgrad(g, 0, g->tM, g->T[X0]));
    return SR_OK;
}

int sr::guess_adam_device(sr_guess *g) {
    g->adam_t += 1;
    const double b1 = g->d.beta1, b2 = g->d.beta2;
    const float lr_t = (float)(g->d.lr * std::sqrt(1.0 - std::pow(b2, g->adam_t)) / (1.0 - std::pow(b1, g->adam_t)));
    for (int l = 0; l < SR_GUESS_LAYERS; l++) {
        launch_adam(wsize(l), g->W[l], g->dW[l], g->mW[l], g->vW[l], lr_t, g->d.beta1, g->d.beta2, g->d.eps, g->s);
        launch_adam(kLayers[l].cout, g->b[l], g->db[l], g->mb[l], g->vb[l], lr_t, g->d.beta1, g->d.beta2, g->d.eps,
                    g->s);
        repack(g, l);
    }
    return SR_OK;
}

// ------------------------------------------------------------------------------------------------
// measurement: one layer's kernels timed with HIP events on the handle's stream
// ------------------------------------------------------------------------------------------------
extern "C" int sr_guess_time_layer(sr_guess *g, int l, int which, int reps, double *avg_ms, double *flops) {
    if (!g || l < 0 || l >= SR_GUESS_LAYERS || which < 0 || which > 2 || reps < 1 || !avg_ms)
        return set_err(SR_ERR_ARG, "bad arguments");
    // the input and output tensors of each layer in the forward pass (guess_forward_device)
    static const int in_t[SR_GUESS_LAYERS] = {X0, A0, X1, A1, X2, A2, X3, A3, U2, U1, U0, G0};
    static const int out_t[SR_GUESS_LAYERS] = {A0, S0, A1, S1, A2, S2, A3, BB, G2, G1, G0, -1};
    const LayerSpec &L = kLayers[l];
    auto launch = [&]() -> int {
        if (which == 0) {
            void *out = out_t[l] >= 0 ? (void *)g->T[out_t[l]] : (void *)g->y;
            return conv(g, l, g->T[in_t[l]], out, l == SR_GUESS_LAYERS - 1 ? CONV_F32_CH0 : CONV_BF16_RELU);
        }
        if (which == 1) return conv_adj(g, l, g->tM, g->tF);
        return wgrad(g, l, g->tM, g->T[in_t[l]]);
    };
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int r = launch();
    if (r == SR_OK) r = launch();
    if (r == SR_OK && hipEventRecord(e0, g->s) != hipSuccess) r = set_err(SR_ERR_HIP, "event record");
    for (int i = 0; i < reps && r == SR_OK; i++) r = launch();
    float ms = 0.f;
    if (r == SR_OK && (hipEventRecord(e1, g->s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                       hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        r = set_err(SR_ERR_HIP, "event timing");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r != SR_OK) return r;
    HIPCHK(hipGetLastError());
    *avg_ms = (double)ms / reps;
    if (flops) *flops = 2.0 * 27.0 * L.cin * L.cout * (double)g->lv[L.level].nvox();
    return SR_OK;
}

// ------------------------------------------------------------------------------------------------
// host-array entry points
// ------------------------------------------------------------------------------------------------
extern "C" int sr_guess_forward(sr_guess *g, const float *x0, float *y) {
    if (!g || !x0 || !y) return set_err(SR_ERR_ARG, "null argument");
    const long n0 = g->lv[0].nvox();
    HIPCHK(hipMemcpyAsync(g->stage, x0, sizeof(float) * n0 * SR_GUESS_IN_CH, hipMemcpyHostToDevice, g->s));
    launch_f32_to_bf16(n0 * SR_GUESS_IN_CH, g->stage, g->T[X0], g->s);
    RETCHK(guess_forward_device(g));
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(y, g->y, sizeof(float) * n0, hipMemcpyDeviceToHost, g->s));
    HIPCHK(hipStreamSynchronize(g->s));
    return SR_OK;
}

extern "C" int sr_guess_backward(sr_guess *g, const float *dy) {
    if (!g || !dy) return set_err(SR_ERR_ARG, "null argument");
    HIPCHK(hipMemcpyAsync(g->dyd, dy, sizeof(float) * g->lv[0].nvox(), hipMemcpyHostToDevice, g->s));
    RETCHK(guess_backward_device(g));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g->s));
    return SR_OK;
}

extern "C" int sr_guess_adam(sr_guess *g) {
    if (!g) return set_err(SR_ERR_ARG, "null argument");
    RETCHK(guess_adam_device(g));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g->s));
    return SR_OK;
}

extern "C" int sr_guess_tensor(sr_guess *g, int which, float *host, long *count) {
    if (!g || which < 0 || which >= 18) return set_err(SR_ERR_ARG, "bad tensor index");
    long n = tensor_elems(g, which);
    if (count) *count = n;
    if (!host) return SR_OK;
    launch_bf16_to_f32(n, g->T[which], g->tF, g->s);
    HIPCHK(hipMemcpyAsync(host, g->tF, sizeof(float) * n, hipMemcpyDeviceToHost, g->s));
    HIPCHK(hipStreamSynchronize(g->s));
    return SR_OK;
}
