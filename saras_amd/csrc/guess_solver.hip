// Solver-facing entry points of the learned initial guess (include/saras_guess.h): the guess for
// the next pressure solve and its online training on the converged solution.
//
// Guess = extrapolation 2 p1 - p2 of the previous two solutions plus a correction predicted by the
// network from the extrapolation's residual r = f - L(2 p1 - p2), the last increment p1 - p2 and
// f (each restricted to the network grid and scaled to unit RMS); the correction is the network
// output times kappa * RMS(r), prolonged to the solver grid.  kappa is the running ratio of the
// RMS of the true correction to the RMS of r, measured during training.  Before a guess is
// returned its relative residual is compared with those of the extrapolation and of the warm start
// p1; the best of the three is used.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>

#include "../../include/saras_guess.h"
#include "mg_state.h"
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

namespace {

constexpr int kRedBlocks = 1024;

struct FineDims {
    int nx, ny, nz;
};

__device__ __forceinline__ long fpad(const FineDims &f, int i, int j, int k) {
    return ((long)(i + 1) * (f.ny + 2) + (j + 1)) * (f.nz + 2) + (k + 1);
}

__device__ __forceinline__ bf16_t to_bf16(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}

__device__ __forceinline__ double w1(int a) { return a == 0 ? 0.5 : 0.25; }

__global__ void __launch_bounds__(256) k_extrap(long n, const double *__restrict__ p1, const double *__restrict__ p2,
                                                double *__restrict__ out) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = p2 ? 2.0 * p1[t] - p2[t] : p1[t];
}

// restriction to the network grid (full weighting over the solver's core nodes, or the node
// itself when both grids coincide) of r, p1 - p2 and f
__global__ void __launch_bounds__(256) k_gather(FineDims f, NetGrid g, int ratio, const double *__restrict__ r,
                                                const double *__restrict__ p1, const double *__restrict__ p2,
                                                const double *__restrict__ rhs, float *__restrict__ st4) {
    long v = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.nvox()) return;
    int K = (int)(v % g.nz), J = (int)((v / g.nz) % g.ny), I = (int)(v / ((long)g.nz * g.ny));
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    if (ratio == 1) {
        long q = fpad(f, I, J, K);
        a0 = r[q];
        a1 = p2 ? p1[q] - p2[q] : 0.0;
        a2 = rhs[q];
    } else {
        for (int a = -1; a <= 1; a++) {
            int i = 2 * I + a;
            if (i < 0 || i >= f.nx) continue;
            for (int b = -1; b <= 1; b++) {
                int j = 2 * J + b;
                if (j < 0 || j >= f.ny) continue;
                for (int c = -1; c <= 1; c++) {
                    int k = 2 * K + c;
                    if (k < 0 || k >= f.nz) continue;
                    double w = w1(a) * w1(b) * w1(c);
                    long q = fpad(f, i, j, k);
                    a0 += w * r[q];
                    if (p2) a1 += w * (p1[q] - p2[q]);
                    a2 += w * rhs[q];
                }
            }
        }
    }
    st4[v * 4 + 0] = (float)a0;
    st4[v * 4 + 1] = (float)a1;
    st4[v * 4 + 2] = (float)a2;
    st4[v * 4 + 3] = 0.0f;
}

// per-channel sums of squares of st4 (fixed block count, fixed order)
__global__ void __launch_bounds__(256) k_sumsq_part(long n, const float *__restrict__ st4, double *__restrict__ part) {
    __shared__ double red[256 * 4];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x)
        for (int c = 0; c < 4; c++) {
            double x = st4[v * 4 + c];
            s[c] += x * x;
        }
    for (int c = 0; c < 4; c++) red[c * 256 + threadIdx.x] = s[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int c = 0; c < 4; c++) red[c * 256 + threadIdx.x] += red[c * 256 + threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x * 256];
}

// the four sums over the block partials: 64 lanes per quantity, then a tree in LDS
__global__ void __launch_bounds__(256) k_sum4_final(const double *__restrict__ part, int nb, double *__restrict__ out) {
    __shared__ double red[256];
    int c = threadIdx.x >> 6, l = threadIdx.x & 63;
    double s = 0.0;
    for (int b = l; b < nb; b += 64) s += part[b * 4 + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 32; w > 0; w >>= 1) {
        if (l < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (l == 0) out[c] = red[threadIdx.x];
}

__global__ void __launch_bounds__(256) k_make_input(long n, const float *__restrict__ st4, const double *__restrict__ ss,
                                                    bf16_t *__restrict__ x0) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * SR_GUESS_IN_CH) return;
    long v = t / SR_GUESS_IN_CH;
    int c = (int)(t % SR_GUESS_IN_CH);
    float val = 0.0f;
    if (c < 3) {
        double rms = sqrt(ss[c] / (double)n);
        val = rms > 0.0 ? (float)(st4[v * 4 + c] / rms) : 0.0f;
    }
    x0[t] = to_bf16(val);
}

__device__ __forceinline__ double prolong_y(const NetGrid &g, int ratio, const float *y, int i, int j, int k) {
    if (ratio == 1) return y[((long)i * g.ny + j) * g.nz + k];
    int ia[2] = {i >> 1, (i + 1) >> 1}, ja[2] = {j >> 1, (j + 1) >> 1}, ka[2] = {k >> 1, (k + 1) >> 1};
    int ni = (i & 1) ? 2 : 1, nj = (j & 1) ? 2 : 1, nk = (k & 1) ? 2 : 1;
    double w = ((i & 1) ? 0.5 : 1.0) * ((j & 1) ? 0.5 : 1.0) * ((k & 1) ? 0.5 : 1.0);
    double s = 0.0;
    for (int a = 0; a < ni; a++)
        for (int b = 0; b < nj; b++)
            for (int c = 0; c < nk; c++) s += y[((long)ia[a] * g.ny + ja[b]) * g.nz + ka[c]];
    return w * s;
}

// out = base + kappa RMS(r) P(y) on the solver's core, base on its ghosts
__global__ void __launch_bounds__(256) k_apply(FineDims f, NetGrid g, int ratio, const double *__restrict__ base,
                                               const float *__restrict__ y, const double *__restrict__ ss, double kappa,
                                               double *__restrict__ out) {
    long total = (long)(f.nx + 2) * (f.ny + 2) * (f.nz + 2);
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int k = (int)(t % (f.nz + 2)) - 1;
    long rr = t / (f.nz + 2);
    int j = (int)(rr % (f.ny + 2)) - 1;
    int i = (int)(rr / (f.ny + 2)) - 1;
    double v = base[t];
    if (i >= 0 && j >= 0 && k >= 0 && i < f.nx && j < f.ny && k < f.nz) {
        double S = kappa * sqrt(ss[0] / (double)g.nvox());
        v += S * prolong_y(g, ratio, y, i, j, k);
    }
    out[t] = v;
}

// d = R(p* - extrap) on the network grid into `target`; partial sums of d^2, d and y (the
// pressure is fixed up to a constant, so target and output are compared with their means removed)
__global__ void __launch_bounds__(256) k_delta(FineDims f, NetGrid g, int ratio, const double *__restrict__ pstar,
                                               const double *__restrict__ ex, const float *__restrict__ y,
                                               float *__restrict__ d, double *__restrict__ part) {
    __shared__ double r0[256], r1[256], r2[256];
    double s2 = 0.0, s1 = 0.0, sy = 0.0;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < g.nvox(); v += (long)gridDim.x * blockDim.x) {
        int K = (int)(v % g.nz), J = (int)((v / g.nz) % g.ny), I = (int)(v / ((long)g.nz * g.ny));
        double a = 0.0;
        if (ratio == 1) {
            long q = fpad(f, I, J, K);
            a = pstar[q] - ex[q];
        } else {
            for (int aa = -1; aa <= 1; aa++) {
                int i = 2 * I + aa;
                if (i < 0 || i >= f.nx) continue;
                for (int b = -1; b <= 1; b++) {
                    int j = 2 * J + b;
                    if (j < 0 || j >= f.ny) continue;
                    for (int c = -1; c <= 1; c++) {
                        int k = 2 * K + c;
                        if (k < 0 || k >= f.nz) continue;
                        long q = fpad(f, i, j, k);
                        a += w1(aa) * w1(b) * w1(c) * (pstar[q] - ex[q]);
                    }
                }
            }
        }
        d[v] = (float)a;
        s2 += a * a;
        s1 += a;
        sy += y[v];
    }
    r0[threadIdx.x] = s2;
    r1[threadIdx.x] = s1;
    r2[threadIdx.x] = sy;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            r0[threadIdx.x] += r0[threadIdx.x + w];
            r1[threadIdx.x] += r1[threadIdx.x + w];
            r2[threadIdx.x] += r2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x * 4 + 0] = r0[0];
        part[blockIdx.x * 4 + 1] = r1[0];
        part[blockIdx.x * 4 + 2] = r2[0];
        part[blockIdx.x * 4 + 3] = 0.0;
    }
}

// centred target t = (d - mean d) / S, centred error e = (y - mean y) - t, loss gradient
// dy = 2 e / n (the gradient of sum e^2 through both means), partial sums of e^2 and t^2
__global__ void __launch_bounds__(256) k_loss(long n, const float *__restrict__ d, const float *__restrict__ y,
                                              const double *__restrict__ ss, const double *__restrict__ sd,
                                              double kappa, float *__restrict__ dy, double *__restrict__ part) {
    __shared__ double r0[256], r1[256];
    const double S = kappa * sqrt(ss[0] / (double)n);
    const double inv = S > 0.0 ? 1.0 / S : 0.0;
    const double dm = sd[1] / (double)n, ym = sd[2] / (double)n;
    double e2 = 0.0, t2 = 0.0;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (long)gridDim.x * blockDim.x) {
        double t = (d[v] - dm) * inv;
        double e = (y[v] - ym) - t;
        dy[v] = (float)(2.0 * e / (double)n);
        e2 += e * e;
        t2 += t * t;
    }
    r0[threadIdx.x] = e2;
    r1[threadIdx.x] = t2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            r0[threadIdx.x] += r0[threadIdx.x + w];
            r1[threadIdx.x] += r1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x * 4 + 0] = r0[0];
        part[blockIdx.x * 4 + 1] = r1[0];
        part[blockIdx.x * 4 + 2] = 0.0;
        part[blockIdx.x * 4 + 3] = 0.0;
    }
}

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ------------------------------------------------------------------------------------------------
// entry points
// ------------------------------------------------------------------------------------------------
extern "C" int sr_guess_predict(sr_guess *g, sr_mg *m, const double *p1, const double *p2, const double *rhs,
                                double *p_out, sr_guess_stats *st) {
    if (!g || !m || !p1 || !rhs || !p_out) return set_err(SR_ERR_ARG, "null argument");
    const Lay &L = m->lv[0].L;
    if (L.nx != g->fine[0] || L.ny != g->fine[1] || L.nz != g->fine[2])
        return set_err(SR_ERR_ARG, "solver grid differs from the guess handle's solver grid");
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = g->s;
    const FineDims f{g->fine[0], g->fine[1], g->fine[2]};
    const NetGrid &n0 = g->lv[0];
    // baselines: the warm start p1 and the extrapolation
    hipLaunchKernelGGL(k_extrap, dim3(nblk(g->fine_pad)), dim3(256), 0, s, g->fine_pad, p1, p2, g->extrap);
    HIPCHK(hipStreamSynchronize(s));
    double rel_ex = 0.0, rel_warm = 0.0, rel_learned = 0.0;
    RETCHK(sr_mg_residual(m, g->extrap, rhs, g->rfield, &rel_ex));
    if (p2) RETCHK(sr_mg_residual(m, p1, rhs, nullptr, &rel_warm));
    else rel_warm = rel_ex;
    // network input, forward pass, correction
    hipLaunchKernelGGL(k_gather, dim3(nblk(n0.nvox())), dim3(256), 0, s, f, n0, g->ratio, g->rfield, p1, p2, rhs,
                       g->stage4);
    hipLaunchKernelGGL(k_sumsq_part, dim3(kRedBlocks), dim3(256), 0, s, n0.nvox(), g->stage4, g->rpart);
    hipLaunchKernelGGL(k_sum4_final, dim3(1), dim3(256), 0, s, g->rpart, kRedBlocks, g->dsc);
    hipLaunchKernelGGL(k_make_input, dim3(nblk(n0.nvox() * SR_GUESS_IN_CH)), dim3(256), 0, s, n0.nvox(), g->stage4,
                       g->dsc, g->T[X0]);
    RETCHK(guess_forward_device(g));
    hipLaunchKernelGGL(k_apply, dim3(nblk(g->fine_pad)), dim3(256), 0, s, f, n0, g->ratio, g->extrap, g->y, g->dsc,
                       g->kappa, p_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    g->have_cache = true;
    int choice = 2;
    if (g->kappa > 0.0) {
        RETCHK(sr_mg_residual(m, p_out, rhs, nullptr, &rel_learned));
    } else {
        rel_learned = rel_ex;   // untrained: the correction is zero, the guess is the extrapolation
        choice = 1;
    }
    if (g->policy == 1) {
        if (g->kappa <= 0.0) RETCHK(sr_mg_residual(m, p_out, rhs, nullptr, &rel_learned));
        choice = 2;
    } else {
        if (choice == 2 && !(rel_learned < rel_ex && rel_learned < rel_warm)) choice = rel_ex <= rel_warm ? 1 : 0;
        if (choice == 1 && rel_warm < rel_ex) choice = 0;
    }
    if (choice != 2) {
        const double *src = choice == 1 ? g->extrap : p1;
        HIPCHK(hipMemcpyAsync(p_out, src, sizeof(double) * g->fine_pad, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    auto t1 = std::chrono::steady_clock::now();
    if (st) {
        st->rel_warm = rel_warm;
        st->rel_extrap = rel_ex;
        st->rel_learned = rel_learned;
        st->choice = choice;
        st->predict_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
    return SR_OK;
}

extern "C" int sr_guess_train(sr_guess *g, const double *p_star, double *rel_loss) {
    if (!g || !p_star) return set_err(SR_ERR_ARG, "null argument");
    if (!g->have_cache) return set_err(SR_ERR_STATE, "train needs the forward pass of a preceding predict");
    hipStream_t s = g->s;
    const FineDims f{g->fine[0], g->fine[1], g->fine[2]};
    const NetGrid &n0 = g->lv[0];
    const long n = n0.nvox();
    // true correction on the network grid and its RMS: calibrates kappa
    hipLaunchKernelGGL(k_delta, dim3(kRedBlocks), dim3(256), 0, s, f, n0, g->ratio, p_star, g->extrap, g->y,
                       g->target, g->rpart);
    hipLaunchKernelGGL(k_sum4_final, dim3(1), dim3(256), 0, s, g->rpart, kRedBlocks, g->dsc + 4);
    HIPCHK(hipMemcpyAsync(g->hsc, g->dsc, sizeof(double) * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const double dmean = g->hsc[5] / (double)n;
    const double rms_r = std::sqrt(g->hsc[0] / (double)n);
    const double rms_d = std::sqrt(std::max(g->hsc[4] / (double)n - dmean * dmean, 0.0));
    if (rms_r > 0.0 && rms_d > 0.0) {
        double k_new = rms_d / rms_r;
        g->kappa = g->trained == 0 ? k_new : 0.9 * g->kappa + 0.1 * k_new;
    }
    if (g->kappa <= 0.0) return set_err(SR_ERR_STATE, "no residual or no correction to learn from");
    hipLaunchKernelGGL(k_loss, dim3(kRedBlocks), dim3(256), 0, s, n, g->target, g->y, g->dsc, g->dsc + 4, g->kappa,
                       g->dyd, g->rpart);
    hipLaunchKernelGGL(k_sum4_final, dim3(1), dim3(256), 0, s, g->rpart, kRedBlocks, g->dsc + 8);
    RETCHK(guess_backward_device(g));
    RETCHK(guess_adam_device(g));
    HIPCHK(hipMemcpyAsync(g->hsc + 8, g->dsc + 8, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    g->trained += 1;
    g->have_cache = false;
    if (rel_loss) *rel_loss = g->hsc[9] > 0.0 ? g->hsc[8] / g->hsc[9] : 0.0;
    return SR_OK;
}

extern "C" int sr_guess_extrapolate(sr_guess *g, const double *p1, const double *p2, double *p_out) {
    if (!g || !p1 || !p_out) return set_err(SR_ERR_ARG, "null argument");
    hipLaunchKernelGGL(k_extrap, dim3(nblk(g->fine_pad)), dim3(256), 0, g->s, g->fine_pad, p1, p2, p_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(g->s));
    return SR_OK;
}

extern "C" int sr_guess_set_policy(sr_guess *g, int policy) {
    if (!g || policy < 0 || policy > 1) return set_err(SR_ERR_ARG, "policy must be 0 or 1");
    g->policy = policy;
    return SR_OK;
}

extern "C" int sr_guess_kappa(const sr_guess *g, double *kappa, int *trained) {
    if (!g) return set_err(SR_ERR_ARG, "null argument");
    if (kappa) *kappa = g->kappa;
    if (trained) *trained = g->trained;
    return SR_OK;
}
