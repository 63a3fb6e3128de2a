// C ABI of the field operators on caller device arrays (include/saras_field.h): derivative
// (lib/field/derivative.cc:121-243), vfield / sfield computeDiff, computeNLin, divergence and
// gradient (vfield.cc:104-320, sfield.cc:86-171), the boundary classes with syncData, and the
// implicit Jacobi solves of hydro_d3 / scalar_d3.
//
// Kernels here work on one field at a time over the reference's padded natural layout (k
// fastest), one thread per entry, k across the lanes of a wavefront so that every stencil load is
// a coalesced row; the fused time step (ts_kernels.hip) keeps its own copies of the same
// expressions.  All are HBM-bound (well under one FP64 flop per byte).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/saras_field.h"
#include "../../include/saras_mg.h"
#include "ghost_waves.h"
#include "xcd.h"
#include "mg_state.h"
#include "ts_kernels.h"

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

struct sr_fgrid {
    sr_fgrid_desc d{};
    int N[3] = {0, 0, 0};
    double h[3] = {0, 0, 0};
    double half = 0.5;                   // calcDerivative2*: factor of the xi^2 part (iScheme)
    FShape shape[4]{};                   // SR_LOC_NODE, SR_LOC_VX, SR_LOC_VY, SR_LOC_VZ
    double *dmet = nullptr;
    Metrics M{};
    hipStream_t s = nullptr;
    double *part = nullptr;              // block maxima of the Jacobi residual
    unsigned long long *dmax = nullptr, *hmax = nullptr;
    hipEvent_t jev = nullptr;            // marks a Jacobi residual's read-back
};

namespace {

__host__ __device__ inline long fsz(const FShape &f) { return (long)(f.n[0] + 2) * (f.n[1] + 2) * (f.n[2] + 2); }

__device__ inline long fix(const FShape &f, int i, int j, int k) {
    return ((long)(i + 1) * (f.n[1] + 2) + (j + 1)) * (f.n[2] + 2) + (k + 1);
}
__device__ inline double mt(const Metrics &M, int d, int node, int q, int i) { return M.m[d][node ? 0 : 1][q][i + 1]; }

// padded entry t -> logical (i, j, k), each in -1..n (32-bit: a padded array has < 2^32 entries)
__device__ inline void unflat(const FShape &f, unsigned t, int &i, int &j, int &k) {
    unsigned n2 = (unsigned)f.n[2] + 2u, n1 = (unsigned)f.n[1] + 2u;
    unsigned r = t / n2;
    k = (int)(t - r * n2) - 1;
    unsigned q = r / n1;
    j = (int)(r - q * n1) - 1;
    i = (int)q - 1;
}
// core entry t -> logical (i, j, k), each in 0..n-1
__device__ inline void unflat_core(const FShape &f, unsigned t, int &i, int &j, int &k) {
    unsigned n2 = (unsigned)f.n[2], n1 = (unsigned)f.n[1];
    unsigned r = t / n2;
    k = (int)(t - r * n2);
    unsigned q = r / n1;
    j = (int)(r - q * n1);
    i = (int)q;
}
// the thread's entry: logical workgroups in XCD order (each XCD a contiguous range of rows)
__device__ inline unsigned entry() { return xcd_block() * 256u + threadIdx.x; }

// calcDerivative1 / 2 along d over the whole padded array: the core range along d gets the
// derivative, the two pad layers along d get 0
__global__ void __launch_bounds__(256) k_fd_deriv(FShape f, Metrics M, const double *__restrict__ F,
                                                  double *__restrict__ out, int order, int d, double ih, double half) {
    const unsigned t = entry();
    if ((long)t >= fsz(f)) return;
    int idx[3];
    unflat(f, t, idx[0], idx[1], idx[2]);
    int r = idx[d];
    if (r < 0 || r >= f.n[d]) {
        out[t] = 0.0;
        return;
    }
    long st = d == 0 ? (long)(f.n[1] + 2) * (f.n[2] + 2) : d == 1 ? (long)(f.n[2] + 2) : 1;
    double fp = F[t + st], fm = F[t - st];
    if (order == 1) {
        double v = (fp - fm) / 2.0;
        v *= ih;
        out[t] = mt(M, d, f.s[d], 1, r) * v;
    } else {
        double ih2 = ih * ih;
        double tmp = (fp - fm) / 2.0 * ih;
        double sec = (fp - 2.0 * F[t] + fm) * ih2;
        out[t] = mt(M, d, f.s[d], 2, r) * tmp + half * mt(M, d, f.s[d], 3, r) * sec;
    }
}

// computeDiff of one field: H(core) += d2xx, += d2yy, += d2zz in that order
__global__ void __launch_bounds__(256) k_fd_diff(FShape f, Metrics M, const double *__restrict__ F,
                                                 double *__restrict__ H, double ih0, double ih1, double ih2,
                                                 double half) {
    const unsigned t = entry();
    if ((long)t >= (long)f.n[0] * f.n[1] * f.n[2]) return;
    int idx[3];
    unflat_core(f, t, idx[0], idx[1], idx[2]);
    long q = fix(f, idx[0], idx[1], idx[2]);
    long st[3] = {(long)(f.n[1] + 2) * (f.n[2] + 2), (long)(f.n[2] + 2), 1};
    double ihh[3] = {ih0, ih1, ih2};
    double acc = H[q];
    double f0 = F[q];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double ih = ihh[d];
        double ih2 = ih * ih;
        double fp = F[q + st[d]], fm = F[q - st[d]];
        double tmp = (fp - fm) / 2.0 * ih;
        double sec = (fp - 2.0 * f0 + fm) * ih2;
        double dd = mt(M, d, f.s[d], 2, idx[d]) * tmp + half * mt(M, d, f.s[d], 3, idx[d]) * sec;
        acc += dd;
    }
    H[q] = acc;
}

// computeNLin of one field at location loc (0 node, 1..3 face of component loc - 1), advected by
// (Vx, Vy, Vz): for d = x, y, z, H(core) -= (sum of the slices of V_d) d_d F / nslices.  Slices of
// V_d onto component c's location: the core index alone when d == c, else {core, d-1, c+1,
// (c+1, d-1)}; onto the nodes: {core, d-1} (field.cc:309-575)
__global__ void __launch_bounds__(256) k_fd_nlin(int loc, FShape f, FShape fx, FShape fy, FShape fz, Metrics M,
                                                 const double *__restrict__ F, const double *__restrict__ Vx,
                                                 const double *__restrict__ Vy, const double *__restrict__ Vz,
                                                 double *__restrict__ H, double ih0, double ih1, double ih2) {
    const unsigned t = entry();
    if ((long)t >= (long)f.n[0] * f.n[1] * f.n[2]) return;
    int i, j, k;
    unflat_core(f, t, i, j, k);
    int idx[3] = {i, j, k};
    long q = fix(f, i, j, k);
    long st[3] = {(long)(f.n[1] + 2) * (f.n[2] + 2), (long)(f.n[2] + 2), 1};
    double ihh[3] = {ih0, ih1, ih2};
    const double *V[3] = {Vx, Vy, Vz};
    const FShape *FS[3] = {&fx, &fy, &fz};
    const int c = loc - 1;
    double acc = H[q];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const FShape &fd = *FS[d];
        const double *Vd = V[d];
        double inter = 0.0;
        int ns;
        if (loc == 0) {
            int o[3] = {i, j, k};
            o[d] -= 1;
            inter += Vd[fix(fd, i, j, k)];
            inter += Vd[fix(fd, o[0], o[1], o[2])];
            ns = 2;
        } else if (d == c) {
            inter += Vd[fix(fd, i, j, k)];
            ns = 1;
        } else {
            int o1[3] = {i, j, k}, o2[3] = {i, j, k}, o3[3] = {i, j, k};
            o1[d] -= 1;
            o2[c] += 1;
            o3[c] += 1;
            o3[d] -= 1;
            inter += Vd[fix(fd, i, j, k)];
            inter += Vd[fix(fd, o1[0], o1[1], o1[2])];
            inter += Vd[fix(fd, o2[0], o2[1], o2[2])];
            inter += Vd[fix(fd, o3[0], o3[1], o3[2])];
            ns = 4;
        }
        double ih = ihh[d];
        double dv = (F[q + st[d]] - F[q - st[d]]) / 2.0;
        dv *= ih;
        double deriv = mt(M, d, f.s[d], 1, idx[d]) * dv;
        // / ns with ns = 1, 2 or 4: the same as * (1 / ns), exactly
        acc -= inter * deriv * (ns == 1 ? 1.0 : ns == 2 ? 0.5 : 0.25);
    }
    H[q] = acc;
}

// gradient component c on its core: G = xi_c(face) (F(+1) - F) / h_c
__global__ void __launch_bounds__(256) k_fd_gradient(int c, FShape fc, FShape fp, Metrics M,
                                                     const double *__restrict__ F, double *__restrict__ G, double hc) {
    const unsigned t = entry();
    if ((long)t >= (long)fc.n[0] * fc.n[1] * fc.n[2]) return;
    int i, j, k;
    unflat_core(fc, t, i, j, k);
    int idx[3] = {i, j, k};
    long sp = c == 0 ? (long)(fp.n[1] + 2) * (fp.n[2] + 2) : c == 1 ? (long)(fp.n[2] + 2) : 1;
    long pq = fix(fp, i, j, k);
    G[fix(fc, i, j, k)] = mt(M, c, 0, 1, idx[c]) * (F[pq + sp] - F[pq]) / hc;
}

// divergence over the whole padded node array: core from the face differences, pads 0
__global__ void __launch_bounds__(256) k_fd_divergence(FShape fp, FShape fx, FShape fy, FShape fz, Metrics M,
                                                       const double *__restrict__ Vx, const double *__restrict__ Vy,
                                                       const double *__restrict__ Vz, double *__restrict__ out,
                                                       double h0, double h1, double h2) {
    const unsigned t = entry();
    if ((long)t >= fsz(fp)) return;
    int i, j, k;
    unflat(fp, t, i, j, k);
    if (i < 0 || j < 0 || k < 0 || i >= fp.n[0] || j >= fp.n[1] || k >= fp.n[2]) {
        out[t] = 0.0;
        return;
    }
    double a = mt(M, 0, 1, 1, i) * (Vx[fix(fx, i, j, k)] - Vx[fix(fx, i - 1, j, k)]) / h0;
    double b = mt(M, 1, 1, 1, j) * (Vy[fix(fy, i, j, k)] - Vy[fix(fy, i, j - 1, k)]) / h1;
    double cz = mt(M, 2, 1, 1, k) * (Vz[fix(fz, i, j, k)] - Vz[fix(fz, i, j, k - 1)]) / h2;
    out[t] = a + b + cz;
}

inline unsigned nblocks(long n) { return (unsigned)((n + 255) / 256); }

int check_loc(const sr_fgrid *g, int loc) {
    if (!g) return set_err(SR_ERR_ARG, "null handle");
    if (loc < SR_LOC_NODE || loc > SR_LOC_VZ) return set_err(SR_ERR_ARG, "location must be SR_LOC_NODE..SR_LOC_VZ");
    return SR_OK;
}

// the boundary conditions as ordered waves: syncData, x walls, y walls, then top and bottom
int bc_waves(const sr_fgrid *g, const FShape &f, const sr_bc *bc, std::vector<GhostWave> &ws) {
    int per[3] = {g->d.xPer, g->d.yPer, g->d.zPer};
    for (int w = 0; w < 6; w++) {
        int kd = bc->kind[w];
        if (kd != SR_BC_PERIODIC && kd != SR_BC_DIRICHLET && kd != SR_BC_NEUMANN && kd != SR_BC_NONE)
            return set_err(SR_ERR_ARG, "boundary kind must be SR_BC_PERIODIC, _DIRICHLET, _NEUMANN or _NONE");
        bool p = per[w / 2] != 0;
        if ((kd == SR_BC_PERIODIC && !p) || (p && kd != SR_BC_PERIODIC && kd != SR_BC_NONE))
            return set_err(SR_ERR_ARG, "periodic walls must match the grid's periodic directions");
    }
    ws.clear();
    ws.push_back(sync_wave(f, per[0], per[1]));
    // wall_op kinds: 0 periodicCC, 1 periodicFC, 2 dirichletCC, 3 dirichletFC, 4 neumannCC, 5 neumannFC
    auto cls = [&](int w) {
        int node = f.s[w / 2];
        if (bc->kind[w] == SR_BC_PERIODIC) return node ? 0 : 1;
        if (bc->kind[w] == SR_BC_DIRICHLET) return node ? 2 : 3;
        return node ? 4 : 5;
    };
    for (int d = 0; d < 2; d++) {
        if (per[d]) continue;
        GhostWave w{};
        for (int q = 2 * d; q <= 2 * d + 1; q++)
            if (bc->kind[q] != SR_BC_NONE) add_op(w, wall_op(f, 0, q, cls(q), bc->value[q]));
        if (w.nops) ws.push_back(w);
    }
    GhostWave w{};
    for (int q : {5, 4})
        if (bc->kind[q] != SR_BC_NONE) add_op(w, wall_op(f, per[2], q, cls(q), bc->value[q]));
    if (w.nops) ws.push_back(w);
    return SR_OK;
}

}  // namespace

// ------------------------------------------------------------------------------------------------
// handle
// ------------------------------------------------------------------------------------------------
static int fgrid_create_impl(const sr_fgrid_desc *desc, sr_fgrid *g) {
    g->d = *desc;
    const sr_fgrid_desc &d = g->d;
    int ind[3] = {d.xInd, d.yInd, d.zInd};
    double L[3] = {d.Lx, d.Ly, d.Lz};
    int gt[3] = {d.xGrid, d.yGrid, d.zGrid};
    double beta[3] = {d.betaX, d.betaY, d.betaZ};
    for (int k = 0; k < 3; k++) {
        if (ind[k] < 2 || ind[k] > 12) return set_err(SR_ERR_ARG, "grid size index out of range 2..12");
        if (L[k] <= 0.0) return set_err(SR_ERR_ARG, "domain length must be positive");
    }
    g->half = d.iScheme == 1 ? 0.5 : 1.0;
    std::vector<double> hmet[3][2][4];
    for (int k = 0; k < 3; k++) {
        int n = (1 << ind[k]) + 1;
        g->N[k] = n;
        g->h[k] = 1.0 / (double)(n - 1);
        for (int face = 0; face < 2; face++) {
            for (int q = 0; q < 4; q++) hmet[k][face][q].assign((size_t)(face ? n + 1 : n + 2), 0.0);
            int r = sr_grid_axis(n, L[k], gt[k], beta[k], face, hmet[k][face][0].data(), hmet[k][face][1].data(),
                                 hmet[k][face][2].data(), hmet[k][face][3].data());
            if (r != SR_OK) return set_err(r, "invalid grid description");
        }
    }
    for (int k = 0; k < 3; k++) {
        g->shape[SR_LOC_NODE].n[k] = g->N[k];
        g->shape[SR_LOC_NODE].s[k] = 1;
    }
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < 3; k++) {
            g->shape[SR_LOC_VX + c].n[k] = c == k ? g->N[k] - 1 : g->N[k];
            g->shape[SR_LOC_VX + c].s[k] = c == k ? 0 : 1;
        }
    HIPCHK(hipStreamCreateWithFlags(&g->s, hipStreamNonBlocking));
    size_t tot = 0;
    for (int k = 0; k < 3; k++)
        for (int face = 0; face < 2; face++)
            for (int q = 0; q < 4; q++) tot += hmet[k][face][q].size();
    HIPCHK(hipMalloc((void **)&g->dmet, sizeof(double) * tot));
    size_t off = 0;
    for (int k = 0; k < 3; k++)
        for (int face = 0; face < 2; face++)
            for (int q = 0; q < 4; q++) {
                const std::vector<double> &v = hmet[k][face][q];
                HIPCHK(hipMemcpy(g->dmet + off, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
                g->M.m[k][face][q] = g->dmet + off;
                off += v.size();
            }
    long nparts = ts_jacobi_res_parts();
    for (int loc = 0; loc < 4; loc++) nparts = std::max(nparts, ts_jacobi_vres_parts(g->shape[loc]));
    HIPCHK(hipMalloc((void **)&g->part, sizeof(double) * nparts));
    HIPCHK(hipEventCreateWithFlags(&g->jev, hipEventDisableTiming));
    HIPCHK(hipMalloc((void **)&g->dmax, sizeof(unsigned long long)));
    HIPCHK(hipHostMalloc((void **)&g->hmax, sizeof(unsigned long long), hipHostMallocDefault));
    return SR_OK;
}

extern "C" int sr_fgrid_create(const sr_fgrid_desc *desc, sr_fgrid **out) {
    if (!desc || !out) return set_err(SR_ERR_ARG, "null argument");
    *out = nullptr;
    sr_fgrid *g = new sr_fgrid();
    int r = fgrid_create_impl(desc, g);
    if (r != SR_OK) {
        std::string msg = sr_last_error();
        sr_fgrid_destroy(g);
        return set_err(r, msg);
    }
    *out = g;
    return SR_OK;
}

extern "C" int sr_fgrid_destroy(sr_fgrid *g) {
    if (!g) return SR_OK;
    if (g->s) (void)hipStreamSynchronize(g->s);
    if (g->dmet) (void)hipFree(g->dmet);
    if (g->jev) (void)hipEventDestroy(g->jev);
    if (g->part) (void)hipFree(g->part);
    if (g->dmax) (void)hipFree(g->dmax);
    if (g->hmax) (void)hipHostFree(g->hmax);
    if (g->s) (void)hipStreamDestroy(g->s);
    delete g;
    return SR_OK;
}

extern "C" int sr_fgrid_dims(const sr_fgrid *g, int loc, int *nx, int *ny, int *nz) {
    RETCHK(check_loc(g, loc));
    if (!nx || !ny || !nz) return set_err(SR_ERR_ARG, "null argument");
    *nx = g->shape[loc].n[0];
    *ny = g->shape[loc].n[1];
    *nz = g->shape[loc].n[2];
    return SR_OK;
}

extern "C" int sr_fgrid_stream(const sr_fgrid *g, void **hip_stream) {
    if (!g || !hip_stream) return set_err(SR_ERR_ARG, "null argument");
    *hip_stream = (void *)g->s;
    return SR_OK;
}

extern "C" int sr_fgrid_sync(sr_fgrid *g) {
    if (!g) return set_err(SR_ERR_ARG, "null handle");
    HIPCHK(hipStreamSynchronize(g->s));
    HIPCHK(hipGetLastError());
    return SR_OK;
}

// ------------------------------------------------------------------------------------------------
// operators
// ------------------------------------------------------------------------------------------------
static int deriv(sr_fgrid *g, int loc, int order, int dim, const double *F, double *out) {
    RETCHK(check_loc(g, loc));
    if (dim < 0 || dim > 2) return set_err(SR_ERR_ARG, "dim must be 0..2");
    if (!F || !out) return set_err(SR_ERR_ARG, "null array");
    if ((const double *)out == F) return set_err(SR_ERR_ARG, "out must not alias F");
    const FShape &f = g->shape[loc];
    // the inverse grid spacings are formed on the host, like derivative's ihx, ihy, ihz (derivative.cc:66-68)
    hipLaunchKernelGGL(k_fd_deriv, dim3(nblocks(fsz(f))), dim3(256), 0, g->s, f, g->M, F, out, order, dim,
                       1.0 / g->h[dim], g->half);
    HIPCHK(hipGetLastError());
    return SR_OK;
}

extern "C" int sr_deriv1(sr_fgrid *g, int loc, int dim, const double *F, double *out) {
    return deriv(g, loc, 1, dim, F, out);
}
extern "C" int sr_deriv2(sr_fgrid *g, int loc, int dim, const double *F, double *out) {
    return deriv(g, loc, 2, dim, F, out);
}

static int diff(sr_fgrid *g, int loc, const double *F, double *H) {
    if (!F || !H) return set_err(SR_ERR_ARG, "null array");
    if ((const double *)H == F) return set_err(SR_ERR_ARG, "H must not alias the field");
    const FShape &f = g->shape[loc];
    long n = (long)f.n[0] * f.n[1] * f.n[2];
    hipLaunchKernelGGL(k_fd_diff, dim3(nblocks(n)), dim3(256), 0, g->s, f, g->M, F, H, 1.0 / g->h[0], 1.0 / g->h[1],
                       1.0 / g->h[2], g->half);
    HIPCHK(hipGetLastError());
    return SR_OK;
}

extern "C" int sr_vf_diff(sr_fgrid *g, const double *Vx, const double *Vy, const double *Vz, double *Hx, double *Hy,
                          double *Hz) {
    if (!g) return set_err(SR_ERR_ARG, "null handle");
    RETCHK(diff(g, SR_LOC_VX, Vx, Hx));
    RETCHK(diff(g, SR_LOC_VY, Vy, Hy));
    RETCHK(diff(g, SR_LOC_VZ, Vz, Hz));
    return SR_OK;
}

extern "C" int sr_sf_diff(sr_fgrid *g, const double *F, double *H) {
    if (!g) return set_err(SR_ERR_ARG, "null handle");
    return diff(g, SR_LOC_NODE, F, H);
}

static int nlin(sr_fgrid *g, int loc, const double *F, const double *const V[3], double *H) {
    if (!F || !H || !V[0] || !V[1] || !V[2]) return set_err(SR_ERR_ARG, "null array");
    if ((const double *)H == F || (const double *)H == V[0] || (const double *)H == V[1] || (const double *)H == V[2])
        return set_err(SR_ERR_ARG, "H must not alias an input");
    const FShape &f = g->shape[loc];
    long n = (long)f.n[0] * f.n[1] * f.n[2];
    hipLaunchKernelGGL(k_fd_nlin, dim3(nblocks(n)), dim3(256), 0, g->s, loc, f, g->shape[SR_LOC_VX],
                       g->shape[SR_LOC_VY], g->shape[SR_LOC_VZ], g->M, F, V[0], V[1], V[2], H, 1.0 / g->h[0],
                       1.0 / g->h[1], 1.0 / g->h[2]);
    HIPCHK(hipGetLastError());
    return SR_OK;
}

extern "C" int sr_vf_nlin(sr_fgrid *g, const double *Ux, const double *Uy, const double *Uz, const double *Vx,
                          const double *Vy, const double *Vz, double *Hx, double *Hy, double *Hz) {
    if (!g) return set_err(SR_ERR_ARG, "null handle");
    // each component's H is written from U_c and the three V_d; H_c must alias none of them,
    // including the other components' inputs that later launches read
    const double *V[3] = {Vx, Vy, Vz};
    const double *U[3] = {Ux, Uy, Uz};
    double *H[3] = {Hx, Hy, Hz};
    for (int c = 0; c < 3; c++)
        for (int e = 0; e < 3; e++)
            if ((const double *)H[c] == U[e] || (const double *)H[c] == V[e])
                return set_err(SR_ERR_ARG, "H must not alias an input");
    for (int c = 0; c < 3; c++) RETCHK(nlin(g, SR_LOC_VX + c, U[c], V, H[c]));
    return SR_OK;
}

extern "C" int sr_sf_nlin(sr_fgrid *g, const double *Vx, const double *Vy, const double *Vz, const double *F,
                          double *H) {
    if (!g) return set_err(SR_ERR_ARG, "null handle");
    const double *V[3] = {Vx, Vy, Vz};
    return nlin(g, SR_LOC_NODE, F, V, H);
}

extern "C" int sr_divergence(sr_fgrid *g, const double *Vx, const double *Vy, const double *Vz, double *div) {
    if (!g) return set_err(SR_ERR_ARG, "null handle");
    if (!Vx || !Vy || !Vz || !div) return set_err(SR_ERR_ARG, "null array");
    if ((const double *)div == Vx || (const double *)div == Vy || (const double *)div == Vz)
        return set_err(SR_ERR_ARG, "div must not alias an input");
    const FShape &fp = g->shape[SR_LOC_NODE];
    hipLaunchKernelGGL(k_fd_divergence, dim3(nblocks(fsz(fp))), dim3(256), 0, g->s, fp, g->shape[SR_LOC_VX],
                       g->shape[SR_LOC_VY], g->shape[SR_LOC_VZ], g->M, Vx, Vy, Vz, div, g->h[0], g->h[1], g->h[2]);
    HIPCHK(hipGetLastError());
    return SR_OK;
}

extern "C" int sr_gradient(sr_fgrid *g, const double *F, double *Gx, double *Gy, double *Gz) {
    if (!g) return set_err(SR_ERR_ARG, "null handle");
    double *G[3] = {Gx, Gy, Gz};
    if (!F || !Gx || !Gy || !Gz) return set_err(SR_ERR_ARG, "null array");
    for (int c = 0; c < 3; c++)
        if ((const double *)G[c] == F) return set_err(SR_ERR_ARG, "the gradient must not alias F");
    for (int c = 0; c < 3; c++) {
        const FShape &fc = g->shape[SR_LOC_VX + c];
        long n = (long)fc.n[0] * fc.n[1] * fc.n[2];
        hipLaunchKernelGGL(k_fd_gradient, dim3(nblocks(n)), dim3(256), 0, g->s, c, fc, g->shape[SR_LOC_NODE], g->M, F,
                           G[c], g->h[c]);
    }
    HIPCHK(hipGetLastError());
    return SR_OK;
}

extern "C" int sr_bc_fill(sr_fgrid *g, int loc, const sr_bc *bc, double *F) {
    RETCHK(check_loc(g, loc));
    if (!bc || !F) return set_err(SR_ERR_ARG, "null argument");
    std::vector<GhostWave> ws;
    RETCHK(bc_waves(g, g->shape[loc], bc, ws));
    for (const GhostWave &w : ws) ts_ghost_wave(g->shape[loc], F, w, g->s);
    HIPCHK(hipGetLastError());
    return SR_OK;
}

extern "C" int sr_jacobi_visc(sr_fgrid *g, int loc, const sr_jacobi_desc *jd, const sr_bc *bc, const double *R,
                              double *F, double *G, int *iters) {
    RETCHK(check_loc(g, loc));
    if (!jd || !bc || !R || !F || !G) return set_err(SR_ERR_ARG, "null argument");
    if (F == G || (const double *)F == R || (const double *)G == R) return set_err(SR_ERR_ARG, "F, G and R must be distinct arrays");
    if (jd->dt <= 0.0 || jd->cnTolerance <= 0.0) return set_err(SR_ERR_ARG, "dt and cnTolerance must be positive");
    if (jd->form == 0 && jd->Re <= 0.0) return set_err(SR_ERR_ARG, "Re must be positive");
    if (jd->form != 0 && jd->form != 1) return set_err(SR_ERR_ARG, "form must be 0 or 1");
    const FShape &f = g->shape[loc];
    std::vector<GhostWave> ws;
    RETCHK(bc_waves(g, f, bc, ws));
    JacobiCoef J = jacobi_coef(g->h, jd->dt, jd->form == 0 ? jd->Re : 0.0);
    J.vis = jd->coef;
    J.form = jd->form;
    for (int k = 0; k < 3; k++) J.ms[k] = f.s[k];
    if (jd->zCollocMetric) J.ms[2] = 0;
    int per[3] = {g->d.xPer, g->d.yPer, g->d.zPer};
    Bulk B = jacobi_bulk(f, per);
    long maxIterations = (long)(g->N[0] - 1) * (g->N[1] - 1) * (g->N[2] - 1);
    double *a = F, *b = G;
    JacobiIO io{g->part, g->dmax, g->hmax, g->jev, g->s};
    int sweeps = 0;
    hipError_t err = hipSuccess;
    int st = ts_jacobi_loop(f, g->M, B, a, b, R, J, ws.data(), (int)ws.size(), io, jd->cnTolerance, maxIterations,
                            &sweeps, &err);
    HIPCHK(err);
    int status = SR_OK;
    if (st == 1) status = set_err(SR_ERR_NOT_CONVERGED, "Jacobi iterations did not converge within the iteration limit");
    // the result lives in whichever buffer holds the converged iterate; the dropped iterate's BCs
    // are still on the stream, so the call waits for them too before G goes back to the caller
    if (a != F) HIPCHK(hipMemcpyAsync(F, a, sizeof(double) * fsz(f), hipMemcpyDeviceToDevice, g->s));
    HIPCHK(hipStreamSynchronize(g->s));
    if (iters) *iters = sweeps;
    return status;
}
