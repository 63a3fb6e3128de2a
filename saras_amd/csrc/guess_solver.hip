// Solver-facing entry points of the learned initial guess (include/saras_guess.h): the guess for
// the next pressure solve and its online training on the converged solution.
//
// Guess = a base guess plus a correction predicted by the network from the base's residual
// r = f - L(base), the last increment p1 - p2 and f (each restricted to the network grid and
// scaled to unit RMS); the correction is the network output times kappa * RMS(r), prolonged to the
// solver grid.  kappa is the running ratio of the RMS of the true correction to the RMS of r,
// measured during training.  The base is the linear predictor when a history is kept (the
// least-squares fit of the new right-hand side by the previous, projected ones, applied to their
// solutions: a Gram-Schmidt basis q_i of right-hand sides with L x_i = q_i, so the fit's residual
// is known without a stencil pass), otherwise the extrapolation 2 p1 - p2.  The returned guess is
// the one with the lowest relative residual.
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

// ------------------------------------------------------------------------------------------------
// linear predictor: projection of the right-hand side, inner products with the basis, linear
// combinations; natural padded layout, core nodes (i, j, k) in [0, n)
// ------------------------------------------------------------------------------------------------
struct Ptr8 {
    const double *p[SR_GUESS_MAX_HISTORY];
};
struct Coef8 {
    double c[SR_GUESS_MAX_HISTORY];
};

__device__ __forceinline__ void core_of(const FineDims &f, unsigned t, int &i, int &j, int &k) {
    unsigned r = t / (unsigned)f.nz;
    k = (int)(t - r * (unsigned)f.nz);
    unsigned i0 = r / (unsigned)f.ny;
    j = (int)(r - i0 * (unsigned)f.ny);
    i = (int)i0;
}

// per-block sums of width W (W <= 16) of a 256-thread block: each wave adds its lanes with
// shuffles, then one thread per column adds the four wave sums in a fixed order; part[b * W + c]
template <int W>
__device__ inline void block_sums(double (&v)[W], double *part) {
    __shared__ double red[W][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < W; c++) {
        double x = v[c];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (lane == 0) red[c][wv] = x;
    }
    __syncthreads();
    if ((int)threadIdx.x < W) {
        const int c = threadIdx.x;
        part[(long)blockIdx.x * W + c] = (red[c][0] + red[c][1]) + (red[c][2] + red[c][3]);
    }
}

// sum over the core of w_x(i) w_y(j) w_z(k) rhs (the left-null-vector weight of the projection)
__global__ void __launch_bounds__(256) k_wsum(FineDims f, const double *__restrict__ rhs, const double *__restrict__ wx,
                                              const double *__restrict__ wy, const double *__restrict__ wz,
                                              double *__restrict__ part) {
    double v[1] = {0.0};
    unsigned n = (unsigned)f.nx * f.ny * f.nz;
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        int i, j, k;
        core_of(f, t, i, j, k);
        v[0] += wx[i] * wy[j] * wz[k] * rhs[fpad(f, i, j, k)];
    }
    block_sums<1>(v, part);
}

// out = rhs - c on the core (c = sums[0] * scale, or 0 without sums), 0 on the ghosts
__global__ void __launch_bounds__(256) k_project_rhs(FineDims f, const double *__restrict__ rhs,
                                                     const double *__restrict__ sums, double scale,
                                                     double *__restrict__ out) {
    unsigned n1 = (unsigned)f.ny + 2, n2 = (unsigned)f.nz + 2;
    unsigned total = ((unsigned)f.nx + 2) * n1 * n2;
    unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    unsigned r = t / n2;
    int k = (int)(t - r * n2) - 1;
    unsigned i0 = r / n1;
    int j = (int)(r - i0 * n1) - 1;
    int i = (int)i0 - 1;
    bool core = i >= 0 && j >= 0 && k >= 0 && i < f.nx && j < f.ny && k < f.nz;
    double c = sums ? sums[0] * scale : 0.0;
    out[t] = core ? rhs[t] - c : 0.0;
}

// (b_q, y) for q < nb and (y, y) over the core: part[b * 9 + q], q = nb holds (y, y)
__global__ void __launch_bounds__(256) k_dots(FineDims f, const double *__restrict__ y, Ptr8 B, int nb,
                                              double *__restrict__ part) {
    double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    unsigned n = (unsigned)f.nx * f.ny * f.nz;
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        int i, j, k;
        core_of(f, t, i, j, k);
        long q = fpad(f, i, j, k);
        double yv = y[q];
#pragma unroll
        for (int a = 0; a < SR_GUESS_MAX_HISTORY; a++)
            if (a < nb) v[a] += B.p[a][q] * yv;
        v[8] += yv * yv;
    }
    if (nb < 8) {   // (y, y) goes to slot nb
        v[nb] = v[8];
        v[8] = 0.0;
    }
    block_sums<9>(v, part);
}

// out = a0 base + sum_q c_q X_q over every padded entry; with part, per-block sums of out^2 over
// the core
__global__ void __launch_bounds__(256) k_lincomb(FineDims f, const double *__restrict__ base, double a0, Ptr8 X,
                                                 Coef8 c, int nx, double *__restrict__ out, double *__restrict__ part) {
    unsigned n1 = (unsigned)f.ny + 2, n2 = (unsigned)f.nz + 2;
    unsigned total = ((unsigned)f.nx + 2) * n1 * n2;
    double v[1] = {0.0};
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        double s = base ? a0 * base[t] : 0.0;
#pragma unroll
        for (int a = 0; a < SR_GUESS_MAX_HISTORY; a++)
            if (a < nx) s += c.c[a] * X.p[a][t];
        out[t] = s;
        if (part) {
            unsigned r = t / n2;
            int k = (int)(t - r * n2) - 1;
            unsigned i0 = r / n1;
            int j = (int)(r - i0 * n1) - 1;
            int i = (int)i0 - 1;
            if (i >= 0 && j >= 0 && k >= 0 && i < f.nx && j < f.ny && k < f.nz) v[0] += s * s;
        }
    }
    if (part) block_sums<1>(v, part);
}

// out[c] = sum over blocks of part[b * W + c], c < W: thread t adds rows t, t + 1024, ... (four
// rows loaded before any is added), then each column is added over the block in a fixed order
// (wave shuffles, then the 16 wave sums by one thread)
template <int W>
__global__ void __launch_bounds__(1024) k_sumw_final(const double *__restrict__ part, int nb,
                                                     double *__restrict__ out) {
    double s[W];
#pragma unroll
    for (int c = 0; c < W; c++) s[c] = 0.0;
    for (int b0 = (int)threadIdx.x; b0 < nb; b0 += 4 * 1024) {
        double v[4][W];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int b = b0 + u * 1024;
#pragma unroll
            for (int c = 0; c < W; c++) v[u][c] = b < nb ? part[(long)b * W + c] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int c = 0; c < W; c++) s[c] += v[u][c];
    }
    __shared__ double red[W][16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < W; c++) {
        double x = s[c];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if (lane == 0) red[c][wv] = x;
    }
    __syncthreads();
    if ((int)threadIdx.x < W) {
        double a = 0.0;
        for (int q = 0; q < 16; q++) a += red[threadIdx.x][q];
        out[threadIdx.x] = a;
    }
}

// Plane-chunk geometry over the padded array: block b covers in-plane entries
// [256 (b % bpp), +256) of planes b / bpp, b / bpp + groups, ...; the in-plane split (j, k) is
// fixed per thread, so the loop over planes needs no division.
struct PlaneChunks {
    unsigned n1, n2, planes, bpp, groups;
    float inv2;
};

__device__ __forceinline__ unsigned fdiv_u(unsigned r, unsigned d, float inv) {
    unsigned q = (unsigned)((float)r * inv);
    int rem = (int)r - (int)(q * d);
    q += (unsigned)(rem >= (int)d) - (unsigned)(rem < 0);
    return q;
}

// k_wsum in plane chunks (in-plane entries below 2^24, where fdiv_u is exact)
__global__ void __launch_bounds__(256) k_wsum_pc(FineDims f, PlaneChunks pc, const double *__restrict__ rhs,
                                                 const double *__restrict__ wx, const double *__restrict__ wy,
                                                 const double *__restrict__ wz, double *__restrict__ part) {
    double v[1] = {0.0};
    const unsigned chunk = blockIdx.x % pc.bpp, g = blockIdx.x / pc.bpp;
    const unsigned r = chunk * 256u + threadIdx.x;
    const unsigned jj = fdiv_u(r, pc.n2, pc.inv2), kk = r - jj * pc.n2;
    const int j = (int)jj - 1, k = (int)kk - 1;
    if (r < pc.n1 * pc.n2 && j >= 0 && j < f.ny && k >= 0 && k < f.nz) {
        const double wjk = wy[j] * wz[k];
        const long ps = (long)pc.n1 * pc.n2;
        // planes 1 + g + q groups in batches of eight loads issued before any is added (plane order)
        const unsigned nq = 1 + g <= (unsigned)f.nx ? ((unsigned)f.nx - 1 - g) / pc.groups + 1 : 0;
        unsigned q = 0;
        for (; q + 8 <= nq; q += 8) {
            double x[8];
#pragma unroll
            for (int u = 0; u < 8; u++) x[u] = rhs[(1 + g + (q + u) * pc.groups) * ps + r];
#pragma unroll
            for (int u = 0; u < 8; u++) v[0] += wx[g + (q + u) * pc.groups] * wjk * x[u];
        }
        for (; q < nq; q++) v[0] += wx[g + q * pc.groups] * wjk * rhs[(1 + g + q * pc.groups) * ps + r];
    }
    block_sums<1>(v, part);
}

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ------------------------------------------------------------------------------------------------
// linear predictor (host side)
// ------------------------------------------------------------------------------------------------
static bool linear_on(const sr_guess *g, const sr_mg *m) {
    return g->hcap > 0 && m->d.semantics == SR_SEM_PROJECTED;
}

static bool singular_mg(const sr_mg *m) { return !(m->d.testPoisson && !(m->d.xPer && m->d.yPer && m->d.zPer)); }

static Ptr8 ptrs(double *const *a, int n) {
    Ptr8 P{};
    for (int i = 0; i < n; i++) P.p[i] = a[i];
    return P;
}

// least squares min ||y - sum c_i f_i|| from the Gram matrix G = (f_i, f_j) and d = (f_i, y):
// c = G^+ d through a Jacobi eigen-decomposition, eigenvalues below 1e-14 of the largest dropped
// (near-dependent right-hand sides); returns ||y - sum c_i f_i||^2 = yy - 2 c.d + c^T G c
static double lsq(int n, const double (*Gm)[SR_GUESS_MAX_HISTORY], const double *d, double yy, double *c) {
    double A[SR_GUESS_MAX_HISTORY][SR_GUESS_MAX_HISTORY], V[SR_GUESS_MAX_HISTORY][SR_GUESS_MAX_HISTORY];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            A[i][j] = Gm[i][j];
            V[i][j] = i == j ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; i++) {
            diag += A[i][i] * A[i][i];
            for (int j = i + 1; j < n; j++) off += A[i][j] * A[i][j];
        }
        if (off <= 1e-32 * diag) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                if (A[p][q] == 0.0) continue;
                double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
                double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < n; k++) {
                    double akp = A[k][p], akq = A[k][q];
                    A[k][p] = cs * akp - sn * akq;
                    A[k][q] = sn * akp + cs * akq;
                }
                for (int k = 0; k < n; k++) {
                    double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = cs * apk - sn * aqk;
                    A[q][k] = sn * apk + cs * aqk;
                }
                for (int k = 0; k < n; k++) {
                    double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = cs * vkp - sn * vkq;
                    V[k][q] = sn * vkp + cs * vkq;
                }
            }
    }
    double lmax = 0.0;
    for (int k = 0; k < n; k++) lmax = std::max(lmax, A[k][k]);
    for (int i = 0; i < n; i++) c[i] = 0.0;
    for (int k = 0; k < n; k++) {
        if (!(A[k][k] > 1e-14 * lmax)) continue;
        double vd = 0.0;
        for (int i = 0; i < n; i++) vd += V[i][k] * d[i];
        for (int i = 0; i < n; i++) c[i] += V[i][k] * vd / A[k][k];
    }
    double r2 = yy;
    for (int i = 0; i < n; i++) {
        r2 -= 2.0 * c[i] * d[i];
        for (int j = 0; j < n; j++) r2 += c[i] * Gm[i][j] * c[j];
    }
    return std::max(r2, 0.0);
}

// fcur = the right-hand side projected like the solver projects it; hd = (f_i, fcur), yy, and the
// least-squares coefficients hc with the fit's residual
static int lin_fit(sr_guess *g, sr_mg *m, const double *rhs) {
    hipStream_t s = g->s;
    const FineDims f{g->fine[0], g->fine[1], g->fine[2]};
    if (singular_mg(m)) {
        PlaneChunks pc;
        pc.n1 = (unsigned)f.ny + 2;
        pc.n2 = (unsigned)f.nz + 2;
        pc.planes = (unsigned)f.nx + 2;
        pc.bpp = (pc.n1 * pc.n2 + 255u) / 256u;
        pc.groups = std::max(1u, std::min((unsigned)f.nx, (unsigned)kLinBlocks / pc.bpp));
        pc.inv2 = 1.0f / (float)pc.n2;
        const unsigned nb = pc.bpp * pc.groups;
        if ((long)pc.n1 * pc.n2 < (1L << 24) && nb <= (unsigned)kLinBlocks * 16u) {
            hipLaunchKernelGGL(k_wsum_pc, dim3(nb), dim3(256), 0, s, f, pc, rhs, m->wvec[0], m->wvec[1], m->wvec[2],
                               g->lpart);
            hipLaunchKernelGGL(k_sumw_final<1>, dim3(1), dim3(1024), 0, s, g->lpart, (int)nb, g->lsum);
        } else {
            hipLaunchKernelGGL(k_wsum, dim3(kLinBlocks), dim3(256), 0, s, f, rhs, m->wvec[0], m->wvec[1],
                               m->wvec[2], g->lpart);
            hipLaunchKernelGGL(k_sumw_final<1>, dim3(1), dim3(1024), 0, s, g->lpart, kLinBlocks, g->lsum);
        }
        hipLaunchKernelGGL(k_project_rhs, dim3(nblk(g->fine_pad)), dim3(256), 0, s, f, rhs, g->lsum, 1.0 / m->wsum,
                           g->fcur);
    } else {
        hipLaunchKernelGGL(k_project_rhs, dim3(nblk(g->fine_pad)), dim3(256), 0, s, f, rhs, (const double *)nullptr,
                           0.0, g->fcur);
    }
    hipLaunchKernelGGL(k_dots, dim3(kLinBlocks), dim3(256), 0, s, f, g->fcur, ptrs(g->hf, g->hn), g->hn, g->lpart);
    hipLaunchKernelGGL(k_sumw_final<9>, dim3(1), dim3(1024), 0, s, g->lpart, kLinBlocks, g->lsum);
    HIPCHK(hipMemcpyAsync(g->lsum_h, g->lsum, sizeof(double) * 9, hipMemcpyDeviceToHost, s));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    for (int i = 0; i < g->hn; i++) g->hd[i] = g->lsum_h[i];
    g->yy = g->lsum_h[g->hn];
    g->fit_r2 = g->hn > 0 ? lsq(g->hn, g->G, g->hd, g->yy, g->hc) : g->yy;
    g->have_fit = true;
    return SR_OK;
}

// relative residual of the fit, ||fcur - sum hc_i f_i|| / ||fcur||
static double lin_rel(const sr_guess *g) { return g->yy > 0.0 ? std::sqrt(g->fit_r2 / g->yy) : 0.0; }

static int lin_apply(sr_guess *g, double *out) {
    const FineDims f{g->fine[0], g->fine[1], g->fine[2]};
    Coef8 c{};
    for (int i = 0; i < g->hn; i++) c.c[i] = g->hc[i];
    hipLaunchKernelGGL(k_lincomb, dim3(kLinBlocks), dim3(256), 0, g->s, f, (const double *)nullptr, 0.0,
                       ptrs(g->hp, g->hn), c, g->hn, out, (double *)nullptr);
    HIPCHK(hipGetLastError());
    return SR_OK;
}

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
    double rel_ex = -1.0, rel_warm = -1.0, rel_learned = -1.0, rel_lin = -1.0;
    // base guess: the linear predictor when it has a history, else the extrapolation
    const double *base = g->extrap;
    int base_choice = 1;
    double rel_base = 0.0;
    bool rfield_ok = false;
    // whether the network's correction is tried on top of the base
    bool use_net = g->policy != 3;
    if (g->policy == 2 && g->net_skip > 0) {
        g->net_skip -= 1;
        use_net = false;
    }
    g->have_fit = false;
    if (linear_on(g, m)) RETCHK(lin_fit(g, m, rhs));
    if (linear_on(g, m) && g->hn >= 1) {
        // without the network the combination is the guess itself: written straight to p_out
        base = use_net ? g->xlin : p_out;
        RETCHK(lin_apply(g, use_net ? g->xlin : p_out));
        rel_lin = lin_rel(g);
        base_choice = 3;
        rel_base = rel_lin;
    } else {
        hipLaunchKernelGGL(k_extrap, dim3(nblk(g->fine_pad)), dim3(256), 0, s, g->fine_pad, p1, p2, g->extrap);
        HIPCHK(hipStreamSynchronize(s));
        RETCHK(sr_mg_residual(m, g->extrap, rhs, g->rfield, &rel_ex));
        rfield_ok = true;
        if (p2) RETCHK(sr_mg_residual(m, p1, rhs, nullptr, &rel_warm));
        rel_base = rel_ex;
    }
    // the network's correction on top of the base
    int choice = base_choice;
    g->have_cache = false;
    if (use_net) {
        HIPCHK(hipStreamSynchronize(s));
        if (!rfield_ok) RETCHK(sr_mg_residual(m, base, rhs, g->rfield, &rel_base));
        if (base_choice == 3) rel_lin = rel_base;   // exact now
        hipLaunchKernelGGL(k_gather, dim3(nblk(n0.nvox())), dim3(256), 0, s, f, n0, g->ratio, g->rfield, p1, p2, rhs,
                           g->stage4);
        hipLaunchKernelGGL(k_sumsq_part, dim3(kRedBlocks), dim3(256), 0, s, n0.nvox(), g->stage4, g->rpart);
        hipLaunchKernelGGL(k_sum4_final, dim3(1), dim3(256), 0, s, g->rpart, kRedBlocks, g->dsc);
        hipLaunchKernelGGL(k_make_input, dim3(nblk(n0.nvox() * SR_GUESS_IN_CH)), dim3(256), 0, s, n0.nvox(),
                           g->stage4, g->dsc, g->T[X0]);
        RETCHK(guess_forward_device(g));
        hipLaunchKernelGGL(k_apply, dim3(nblk(g->fine_pad)), dim3(256), 0, s, f, n0, g->ratio, base, g->y, g->dsc,
                           g->kappa, p_out);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
        g->have_cache = true;
        g->train_base = base;
        if (g->kappa > 0.0 || g->policy == 1) {
            RETCHK(sr_mg_residual(m, p_out, rhs, nullptr, &rel_learned));
            if (g->policy == 1 || rel_learned < rel_base) choice = 2;
        }
        if (g->policy == 0 && g->kappa > 0.0) {
            // keep the record while training so that the adaptive policy starts from it
            g->net_losses = choice == 2 ? 0 : g->net_losses + 1;
        } else if (g->policy == 2) {
            if (choice == 2) {
                g->net_losses = 0;
                g->net_backoff = 8;
            } else if (++g->net_losses >= 3) {
                // skip for a while, twice as long after every further losing probe (at most 256)
                g->net_skip = g->net_backoff;
                g->net_backoff = std::min(2 * g->net_backoff, 256);
                g->net_losses = 2;
            }
        }
    }
    // without a history the warm start competes too
    if (rel_warm >= 0.0 && g->policy != 1) {
        double best = choice == 2 ? rel_learned : rel_ex;
        if (rel_warm < best) choice = 0;
    }
    if (choice != 2) {
        const double *src = choice == 0 ? p1 : base;
        if (src != p_out) HIPCHK(hipMemcpyAsync(p_out, src, sizeof(double) * g->fine_pad, hipMemcpyDeviceToDevice, s));
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    auto t1 = std::chrono::steady_clock::now();
    if (st) {
        st->rel_warm = rel_warm;
        st->rel_extrap = rel_ex;
        st->rel_learned = rel_learned;
        st->rel_linear = rel_lin;
        st->choice = choice;
        st->net_used = use_net ? 1 : 0;
        st->predict_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
    return SR_OK;
}

extern "C" int sr_guess_observe(sr_guess *g, sr_mg *m, const double *rhs, const double *p_star) {
    if (!g || !m || !rhs || !p_star) return set_err(SR_ERR_ARG, "null argument");
    if (!linear_on(g, m)) return SR_OK;
    const Lay &L = m->lv[0].L;
    if (L.nx != g->fine[0] || L.ny != g->fine[1] || L.nz != g->fine[2])
        return set_err(SR_ERR_ARG, "solver grid differs from the guess handle's solver grid");
    if (!g->have_fit) RETCHK(lin_fit(g, m, rhs));
    g->have_fit = false;
    if (!(g->yy > 0.0)) return SR_OK;   // a zero right-hand side adds nothing
    // the pair goes to the next ring slot (the oldest once the ring is full)
    int slot = g->hnext;
    hipStream_t s = g->s;
    // the projected right-hand side changes places with the slot's buffer (the oldest pair's, now
    // free), the solution is copied
    std::swap(g->hf[slot], g->fcur);
    HIPCHK(hipMemcpyAsync(g->hp[slot], p_star, sizeof(double) * g->fine_pad, hipMemcpyDeviceToDevice, s));
    int n_new = std::min(g->hn + 1, g->hcap);
    for (int i = 0; i < n_new; i++) {
        if (i == slot) continue;
        g->G[slot][i] = g->G[i][slot] = g->hd[i];
    }
    g->G[slot][slot] = g->yy;
    g->hn = n_new;
    g->hnext = (slot + 1) % g->hcap;
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    return SR_OK;
}

extern "C" int sr_guess_train(sr_guess *g, const double *p_star, double *rel_loss) {
    if (!g || !p_star) return set_err(SR_ERR_ARG, "null argument");
    if (!g->have_cache) {   // the last predict skipped the network: nothing to learn from
        if (rel_loss) *rel_loss = -1.0;
        return SR_OK;
    }
    hipStream_t s = g->s;
    const FineDims f{g->fine[0], g->fine[1], g->fine[2]};
    const NetGrid &n0 = g->lv[0];
    const long n = n0.nvox();
    // true correction on the network grid and its RMS: calibrates kappa
    hipLaunchKernelGGL(k_delta, dim3(kRedBlocks), dim3(256), 0, s, f, n0, g->ratio, p_star, g->train_base, g->y,
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
    if (!g || policy < 0 || policy > 3) return set_err(SR_ERR_ARG, "policy must be 0..3");
    g->policy = policy;
    g->net_skip = 0;
    g->net_backoff = 8;
    // a network that lost its last three comparisons starts the adaptive policy skipped
    if (policy == 2 && g->net_losses >= 3) {
        g->net_skip = g->net_backoff;
        g->net_backoff *= 2;
        g->net_losses = 2;
    }
    return SR_OK;
}

extern "C" int sr_guess_kappa(const sr_guess *g, double *kappa, int *trained) {
    if (!g) return set_err(SR_ERR_ARG, "null argument");
    if (kappa) *kappa = g->kappa;
    if (trained) *trained = g->trained;
    return SR_OK;
}
