// C ABI of the multigrid solver (include/saras_mg.h): level hierarchy, V-cycle orchestration,
// hipGraph capture of the cycle, and the per-operator entry points.
//
// Restates the control flow of lib/poisson/poisson.cc (mgSolve 100-151, vCycle 219-280) and the
// construction of lib/poisson/poisson_d3.cc:64-87 on the device kernels of mg_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/saras_mg.h"
#include "mg_kernels.h"
#include "mg_state.h"

using namespace sr;

// ------------------------------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------------------------------
static thread_local std::string g_err;

extern "C" const char *sr_last_error(void) { return g_err.c_str(); }

int sr::set_err(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

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

// ------------------------------------------------------------------------------------------------
// construction
// ------------------------------------------------------------------------------------------------
static Lay make_lay(int nx, int ny, int nz) {
    Lay L;
    L.nx = nx;
    L.ny = ny;
    L.nz = nz;
    L.kh = (nz + 1) / 2 + 1;
    L.khs = L.kh;
    return L;
}

// Left null vector of one direction (periodic: trapezoid weights; mirror ghosts: the balance
// w_k c_k = w_{k+1} a_{k+1} of a zero-row-sum three-point operator).
static std::vector<double> null_vector(int n, bool periodic, double h, const double *d2, const double *d1) {
    std::vector<double> w(n, 1.0);
    if (periodic) {
        w[0] = 0.5;
        w[n - 1] = 0.5;
        return w;
    }
    double h2 = h * h;
    w[0] = 1.0;
    for (int k = 0; k < n - 1; k++) {
        double ck = d2[k] / h2 + d1[k] / (2.0 * h);
        double ak1 = d2[k + 1] / h2 - d1[k + 1] / (2.0 * h);
        if (k == 0) ck += d2[0] / h2 - d1[0] / (2.0 * h);
        if (k + 1 == n - 1) ak1 += d2[n - 1] / h2 + d1[n - 1] / (2.0 * h);
        w[k + 1] = w[k] * ck / ak1;
    }
    return w;
}

template <typename T>
static int upload_vec(T **dst, const std::vector<T> &v) {
    HIPCHK(hipMalloc((void **)dst, sizeof(T) * std::max<size_t>(v.size(), 1)));
    HIPCHK(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return SR_OK;
}

// Coarse end of the V-cycle in one workgroup: the largest level index range first..depth whose u, f
// and coefficients fit the workgroup's LDS (first >= 1: level 0 carries the solution itself).
template <typename T>
static void plan_coarse(sr_mg *m, int &first, CoarseDesc &D, size_t &lds, size_t limit) {
    first = 1 << 20;
    int depth = m->depth;
    for (int f = 1; f < depth; f++) {
        CoarseDesc t{};
        t.first = f;
        t.last = depth;
        if (depth - f + 1 > 8) continue;
        for (int q = 0; q <= depth - f; q++) t.L[q] = m->lv[f + q].L;
        size_t bytes = coarse_cycle_plan<T>(t, limit);
        if (bytes <= limit) {
            first = f;
            D = t;
            lds = bytes;
            break;
        }
    }
}

// Dense inverse of the coarsest-level operator for coarseDirect.  A (nc x nc) is the 7-point
// operator of the coarsest level with its face-ghost images folded in (periodic wrap, mirror,
// antisymmetric image; walls homogeneous, as on every coarse level).  A singular A is bordered with
// the constant vector, B = [[A, 1], [1^T, 0]]: for any f, B^{-1} [f; 0] = [u; lambda] with
// A u = f - lambda 1 and sum u = 0, i.e. the solution of the RHS projected along the constant
// vector (lambda = 0 for a consistent f).  The nc x nc block of B^{-1} is stored transposed.
static bool coarse_singular(const sr_mg *m) { return !(m->d.testPoisson && !(m->d.xPer && m->d.yPer && m->d.zPer)); }

static int build_coarse_direct(sr_mg *m) {
    const Level &Cv = m->lv[m->depth];
    const Lay &L = Cv.L;
    int n[3] = {L.nx, L.ny, L.nz};
    int nc = n[0] * n[1] * n[2];
    if (nc > 1024) return set_err(SR_ERR_ARG, "coarseDirect needs at most 1024 coarsest-level nodes");
    Bc bc = pressure_bc(m, m->depth, 1);
    bool sing = coarse_singular(m);
    int N = sing ? nc + 1 : nc;
    std::vector<double> B((size_t)N * N, 0.0);
    auto node = [&](const int *p) { return ((p[0] - 1) * n[1] + (p[1] - 1)) * n[2] + (p[2] - 1); };
    for (int i = 1; i <= n[0]; i++)
        for (int j = 1; j <= n[1]; j++)
            for (int k = 1; k <= n[2]; k++) {
                int p[3] = {i, j, k};
                int row = node(p);
                for (int d = 0; d < 3; d++) {
                    const double *c = &Cv.coef_h[d][4 * (size_t)p[d]];
                    B[(size_t)row * N + row] -= c[2];
                    for (int sgn = -1; sgn <= 1; sgn += 2) {
                        double w = sgn > 0 ? c[0] : c[1];
                        int q[3] = {p[0], p[1], p[2]};
                        q[d] += sgn;
                        double mult = 1.0;
                        if (q[d] < 1 || q[d] > n[d]) {
                            int r = bc.rule[d];
                            if (r != G_PERIODIC && r != G_MIRROR && r != G_ANTI) continue;   // homogeneous wall
                            q[d] = ghost_src(r, q[d], n[d]);
                            if (r == G_ANTI) mult = -1.0;
                        }
                        B[(size_t)row * N + node(q)] += mult * w;
                    }
                }
                if (sing) {
                    B[(size_t)row * N + nc] = 1.0;
                    B[(size_t)nc * N + row] = 1.0;
                }
            }
    // Gauss-Jordan inverse with partial pivoting
    std::vector<double> I((size_t)N * N, 0.0);
    for (int r = 0; r < N; r++) I[(size_t)r * N + r] = 1.0;
    for (int col = 0; col < N; col++) {
        int piv = col;
        for (int r = col + 1; r < N; r++)
            if (std::fabs(B[(size_t)r * N + col]) > std::fabs(B[(size_t)piv * N + col])) piv = r;
        if (std::fabs(B[(size_t)piv * N + col]) < 1e-300) return set_err(SR_ERR_ARG, "coarsest operator is singular");
        if (piv != col)
            for (int c = 0; c < N; c++) {
                std::swap(B[(size_t)piv * N + c], B[(size_t)col * N + c]);
                std::swap(I[(size_t)piv * N + c], I[(size_t)col * N + c]);
            }
        double inv = 1.0 / B[(size_t)col * N + col];
        for (int c = 0; c < N; c++) {
            B[(size_t)col * N + c] *= inv;
            I[(size_t)col * N + c] *= inv;
        }
        for (int r = 0; r < N; r++) {
            if (r == col) continue;
            double fct = B[(size_t)r * N + col];
            if (fct == 0.0) continue;
            for (int c = 0; c < N; c++) {
                B[(size_t)r * N + c] -= fct * B[(size_t)col * N + c];
                I[(size_t)r * N + c] -= fct * I[(size_t)col * N + c];
            }
        }
    }
    std::vector<double> pt((size_t)nc * nc);
    for (int i = 0; i < nc; i++)
        for (int j = 0; j < nc; j++) pt[(size_t)j * nc + i] = I[(size_t)i * N + j];
    RETCHK(upload_vec(&m->pinv64, pt));
    std::vector<float> pf(pt.begin(), pt.end());
    RETCHK(upload_vec(&m->pinv32, pf));
    m->ncd = nc;
    return SR_OK;
}

static void finish_coarse(sr_mg *m, CoarseDesc &D, bool fp32) {
    for (int q = 0; q <= D.last - D.first; q++)
        for (int a = 0; a < 3; a++)
            D.cg[q][a] = fp32 ? (const void *)m->lv[D.first + q].c32[a] : (const void *)m->lv[D.first + q].c64[a];
    D.pre = fp32 ? m->pre32 : m->d.preSmooth;
    D.post = fp32 ? m->post32 : m->d.postSmooth;
    D.bc = pressure_bc(m, D.first, 1);
    for (int q = 0; q < 6; q++) D.bc.wall[q] = nullptr;
    D.tbc = tmp_bc(m);
    D.tol = m->d.mgTolerance;
    D.max_count = m->maxCount;
    bool sing = !(m->d.testPoisson && !(m->d.xPer && m->d.yPer && m->d.zPer));
    D.project = (m->d.coarseProject && sing) ? 1 : 0;
    Level &Cv = m->lv[m->depth];
    for (int a = 0; a < 3; a++) D.wc[a] = Cv.wc[a];
    D.wcsum = Cv.wcsum;
    D.pinv = fp32 ? (const void *)m->pinv32 : (const void *)m->pinv64;
    D.ncd = m->ncd;
}

static int create_impl(const sr_mg_desc *d, sr_mg *m) {
    m->d = *d;
    for (int q = 0; q < 6; q++) m->d.wall[q] = nullptr;
    int ind[3] = {d->xInd, d->yInd, d->zInd};
    int n0[3];
    for (int a = 0; a < 3; a++) {
        if (ind[a] < 1 || ind[a] > 14) return set_err(SR_ERR_ARG, "size index out of range 1..14");
        n0[a] = (1 << ind[a]) + 1;
    }
    int depth = d->vcDepth;
    if (depth < 1 || depth >= *std::min_element(ind, ind + 3))
        return set_err(SR_ERR_ARG, "vcDepth must satisfy 1 <= vcDepth < min(xInd, yInd, zInd)");
    if (d->preSmooth < 0 || d->postSmooth < 0 || d->vcCount < 0) return set_err(SR_ERR_ARG, "negative counts");
    m->pre32 = d->innerPreSmooth > 0 ? d->innerPreSmooth : d->preSmooth;
    m->post32 = d->innerPostSmooth > 0 ? d->innerPostSmooth : d->postSmooth;
    const double *met[6] = {d->xix2, d->xixx, d->ety2, d->etyy, d->ztz2, d->ztzz};
    for (int q = 0; q < 6; q++)
        if (!met[q]) return set_err(SR_ERR_ARG, "metric array missing");
    for (int q = 0; q < 6; q++) m->met0[q].assign(met[q], met[q] + n0[q / 2] + 2);

    HIPCHK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&m->ev0));
    HIPCHK(hipEventCreate(&m->ev1));
    HIPCHK(hipEventCreate(&m->sw0));
    HIPCHK(hipEventCreate(&m->sw1));
    m->depth = depth;
    m->lv.resize(depth + 1);
    double dxi[3] = {d->dXi, d->dEt, d->dZt};
    for (int l = 0; l <= depth; l++) {
        Level &V = m->lv[l];
        V.L = make_lay((1 << (ind[0] - l)) + 1, (1 << (ind[1] - l)) + 1, (1 << (ind[2] - l)) + 1);
        int nl[3] = {V.L.nx, V.L.ny, V.L.nz};
        double s = (double)(1 << l);
        for (int a = 0; a < 3; a++) {
            double h = s * dxi[a];   // hx(l) = strideValues(l) * dXi (poisson.cc:420-424)
            V.h[a] = h;
            // coarse metrics = fine metrics sampled at stride 2^l, pads 0 (poisson.cc:467-494)
            std::vector<double> c(4 * (size_t)(nl[a] + 2), 0.0);
            for (int i = 0; i < nl[a]; i++) {
                double m2 = met[2 * a][i * (1 << l) + 1], m1 = met[2 * a + 1][i * (1 << l) + 1];
                c[4 * (i + 1) + 0] = m2 / (h * h) + m1 / (2.0 * h);
                c[4 * (i + 1) + 1] = m2 / (h * h) - m1 / (2.0 * h);
                // the centre coefficient 2 m2 / h^2 as the sum of the two neighbour coefficients (equal
                // to within one rounding); the FP64 kernels form the diagonal from those two instead of
                // loading a third value
                c[4 * (i + 1) + 2] = c[4 * (i + 1) + 0] + c[4 * (i + 1) + 1];
            }
            V.coef_h[a] = c;
            RETCHK(upload_vec(&V.c64[a], c));
            std::vector<float> cf(c.begin(), c.end());
            RETCHK(upload_vec(&V.c32[a], cf));
        }
        long sz = lay_size(V.L);
        HIPCHK(hipMalloc((void **)&V.u, sizeof(double) * sz));
        HIPCHK(hipMalloc((void **)&V.f, sizeof(double) * sz));
        HIPCHK(hipMalloc((void **)&V.t, sizeof(double) * sz));
        launch_fill<double>(V.u, sz, 0.0, m->stream);
        launch_fill<double>(V.f, sz, 0.0, m->stream);
        launch_fill<double>(V.t, sz, 0.0, m->stream);
    }
    const Lay &C = m->lv[depth].L;
    m->maxCount = (C.nx + 1) * (C.ny + 1) * (C.nz + 1);   // poisson.cc:359-361
    if (coarsest_lds_bytes<double>(C) > 60 * 1024)
        return set_err(SR_ERR_ARG, "coarsest level too large for the single-workgroup solve; increase vcDepth");

    // Dirichlet wall planes (TEST_POISSON), level 0 only
    const Lay &F = m->lv[0].L;
    long wsz[3] = {(long)(F.ny + 2) * (F.nz + 2), (long)(F.nx + 2) * (F.nz + 2), (long)(F.nx + 2) * (F.ny + 2)};
    for (int q = 0; q < 6; q++) {
        if (d->wall[q]) {
            std::vector<double> w(d->wall[q], d->wall[q] + wsz[q / 2]);
            RETCHK(upload_vec(&m->wall[q], w));
        }
    }
    if (d->precision != SR_PREC_FP64 && d->precision != SR_PREC_MIXED && d->precision != SR_PREC_MIXED_GCR)
        return set_err(SR_ERR_ARG, "unknown precision mode");
    if (d->precision != SR_PREC_FP64) RETCHK(ensure_fp32(m));
    if (d->precision == SR_PREC_MIXED_GCR) {
        long sz = lay_size(F);
        for (int q = 0; q < 2; q++) HIPCHK(hipMalloc((void **)&m->zdir[q], sizeof(float) * sz));
        HIPCHK(hipMalloc((void **)&m->gpart, sizeof(double) * 5 * (size_t)norm_res_parts(F)));
        HIPCHK(hipMalloc((void **)&m->gcoef, sizeof(double) * 8));
        // wall ghosts of the directions are never written: they stay zero like the correction's
        for (int q = 0; q < 2; q++) launch_fill<float>(m->zdir[q], sz, 0.0f, m->stream);
    }
    if (d->coarseDirect) RETCHK(build_coarse_direct(m));
    // coarse end of the cycle in LDS
    {
        int dev = 0, lds_max = 65536;
        HIPCHK(hipGetDevice(&dev));
        if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) lds_max = 65536;
        m->lds_limit = lds_max;
        size_t limit = (size_t)lds_max - 1024;   // static LDS of the kernel
        if (d->gsSmooth) {
            plan_coarse<double>(m, m->coarse_first64, m->cdesc64, m->clds64, limit);
            if (m->coarse_first64 <= depth) {
                finish_coarse(m, m->cdesc64, false);
                if (coarse_cycle_attr<double>(m->clds64)) m->coarse_first64 = 1 << 20;
            }
            plan_coarse<float>(m, m->coarse_first32, m->cdesc32, m->clds32, limit);
            if (m->coarse_first32 <= depth) {
                finish_coarse(m, m->cdesc32, true);
                if (coarse_cycle_attr<float>(m->clds32)) m->coarse_first32 = 1 << 20;
            }
        }
        (void)hipGetLastError();
    }
    // null vectors for the range projection (finest level)
    const Level &V0 = m->lv[0];
    int per[3] = {d->xPer, d->yPer, d->zPer};
    double wsum[3];
    for (int a = 0; a < 3; a++) {
        int n = n0[a];
        auto w = null_vector(n, per[a] != 0, V0.h[a], met[2 * a] + 1, met[2 * a + 1] + 1);
        wsum[a] = 0;
        for (double x : w) wsum[a] += x;
        RETCHK(upload_vec(&m->wvec[a], w));
    }
    m->wsum = wsum[0] * wsum[1] * wsum[2];
    // null vectors of the coarsest level, from its sampled metrics (coarseProject)
    {
        Level &Cv = m->lv[depth];
        int nc[3] = {Cv.L.nx, Cv.L.ny, Cv.L.nz};
        double prod = 1.0;
        for (int a = 0; a < 3; a++) {
            std::vector<double> d2(nc[a]), d1(nc[a]);
            for (int i = 0; i < nc[a]; i++) {
                d2[i] = met[2 * a][i * (1 << depth) + 1];
                d1[i] = met[2 * a + 1][i * (1 << depth) + 1];
            }
            auto w = null_vector(nc[a], per[a] != 0, Cv.h[a], d2.data(), d1.data());
            double sw = 0;
            for (double x : w) sw += x;
            prod *= sw;
            RETCHK(upload_vec(&Cv.wc[a], w));
        }
        Cv.wcsum = prod;
    }

    int nparts = norm_parts(F);
    HIPCHK(hipMalloc((void **)&m->part, sizeof(double) * 4 * (size_t)nparts));
    HIPCHK(hipMalloc((void **)&m->part2, sizeof(double) * 2 * (size_t)norm_res_parts(F)));
    HIPCHK(hipMalloc((void **)&m->fterms, sizeof(double) * 4));
    HIPCHK(hipMalloc((void **)&m->fstage, sizeof(double) * FINAL_STAGE));
    HIPCHK(hipMalloc((void **)&m->fcount, sizeof(unsigned)));
    HIPCHK(hipMemsetAsync(m->fcount, 0, sizeof(unsigned), m->stream));
    m->fb_coef = FinalBufs{m->fstage, m->fcount};
    launch_fill<double>(m->fterms, 4, 0.0, m->stream);
    HIPCHK(hipMalloc((void **)&m->dres, sizeof(double) * 16));
    HIPCHK(hipMalloc((void **)&m->diters, sizeof(int) * 4));
    HIPCHK(hipHostMalloc((void **)&m->hres, sizeof(double) * 16, hipHostMallocDefault));
    HIPCHK(hipStreamSynchronize(m->stream));
    return SR_OK;
}

extern "C" int sr_mg_create(const sr_mg_desc *desc, sr_mg **out) {
    if (!desc || !out) return set_err(SR_ERR_ARG, "null argument");
    *out = nullptr;
    sr_mg *m = new sr_mg();
    int r = create_impl(desc, m);
    if (r != SR_OK) {
        std::string keep = g_err;
        sr_mg_destroy(m);
        g_err = keep;
        return r;
    }
    *out = m;
    return SR_OK;
}

extern "C" int sr_mg_destroy(sr_mg *m) {
    if (!m) return SR_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    if (m->graph_exec) (void)hipGraphExecDestroy(m->graph_exec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    for (int q = 0; q < 2; q++) {
        if (m->graph_mixed[q]) (void)hipGraphExecDestroy(m->graph_mixed[q]);
        if (m->graph_mixed_g[q]) (void)hipGraphDestroy(m->graph_mixed_g[q]);
        if (m->graph_gcr[q]) (void)hipGraphExecDestroy(m->graph_gcr[q]);
        if (m->graph_gcr_g[q]) (void)hipGraphDestroy(m->graph_gcr_g[q]);
        (void)hipFree(m->zdir[q]);
    }
    (void)hipFree(m->pinv64); (void)hipFree(m->pinv32);
    (void)hipFree(m->gpart); (void)hipFree(m->gcoef);
    for (Level &V : m->lv) {
        (void)hipFree(V.u); (void)hipFree(V.f); (void)hipFree(V.t);
        (void)hipFree(V.u32); (void)hipFree(V.f32); (void)hipFree(V.t32);
        for (int a = 0; a < 3; a++) { (void)hipFree(V.c64[a]); (void)hipFree(V.c32[a]); (void)hipFree(V.wc[a]); }
    }
    for (int q = 0; q < 6; q++) (void)hipFree(m->wall[q]);
    for (int a = 0; a < 3; a++) (void)hipFree(m->wvec[a]);
    (void)hipFree(m->part); (void)hipFree(m->part2); (void)hipFree(m->fterms); (void)hipFree(m->fstage); (void)hipFree(m->fcount); (void)hipFree(m->dres); (void)hipFree(m->diters); (void)hipFree(m->stage);
    (void)hipFree(m->stage2); (void)hipFree(m->r32);
    if (m->hres) (void)hipHostFree(m->hres);
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    if (m->sw0) (void)hipEventDestroy(m->sw0);
    if (m->sw1) (void)hipEventDestroy(m->sw1);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return SR_OK;
}

// ------------------------------------------------------------------------------------------------
// boundary rules
// ------------------------------------------------------------------------------------------------
// imposeBC rules of pressureData at a level (poisson_d3.cc:575-668)
Bc sr::pressure_bc(const sr_mg *m, int level, int zeroBC) {
    Bc b{};
    int per[3] = {m->d.xPer, m->d.yPer, m->d.zPer};
    for (int a = 0; a < 3; a++) {
        if (per[a]) b.rule[a] = G_PERIODIC;
        else if (m->d.testPoisson) b.rule[a] = zeroBC ? G_ANTI : G_WALL;
        else b.rule[a] = G_MIRROR;
    }
    for (int q = 0; q < 6; q++) b.wall[q] = (level == 0) ? m->wall[q] : nullptr;
    return b;
}

// updatePads rules (poisson_d3.cc:671-686): periodic x/y exchange, nothing else touched
static Bc pads_bc(const sr_mg *m) {
    Bc b{};
    b.rule[0] = m->d.xPer ? G_PERIODIC : G_KEEP;
    b.rule[1] = m->d.yPer ? G_PERIODIC : G_KEEP;
    b.rule[2] = G_KEEP;
    return b;
}

// residual pads read by coarsen(): reference = zero except the periodic x/y exchange
// (poisson_d3.cc:91, 110); padMode 1 = follow each direction's ghost rule
Bc sr::tmp_bc(const sr_mg *m) {
    Bc b{};
    int per[3] = {m->d.xPer, m->d.yPer, m->d.zPer};
    for (int a = 0; a < 3; a++) {
        if (m->d.padMode == 1) {
            if (per[a]) b.rule[a] = G_PERIODIC;
            else b.rule[a] = m->d.testPoisson ? G_ZERO : G_MIRROR;
        } else {
            b.rule[a] = (a < 2 && per[a]) ? G_PERIODIC : G_ZERO;
        }
    }
    return b;
}

static bool singular(const sr_mg *m) {
    return !(m->d.testPoisson && !(m->d.xPer && m->d.yPer && m->d.zPer));
}

// ------------------------------------------------------------------------------------------------
// operators (device, FP64 hierarchy)
// ------------------------------------------------------------------------------------------------
// Timing marks inside a captured cycle must be external event nodes; a runtime that refuses them
// loses the in-solve timing, not the solve.
void sr::mark(sr_mg *m, hipEvent_t e) {
    if (!e || m->marks_off) return;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(m->stream, &cs);
    hipError_t r = cs == hipStreamCaptureStatusActive ? hipEventRecordWithFlags(e, m->stream, hipEventRecordExternal)
                                                      : hipEventRecord(e, m->stream);
    if (r != hipSuccess) {
        (void)hipGetLastError();
        m->marks_off = true;
    }
}
// smooth(): ghosts, then `count` red-black sweeps (each updated node refreshes its face ghosts);
// exact = 1 adds the closing imposeBC of poisson_d3.cc:181 (edge and corner ghosts included).
void sr::smooth_level(sr_mg *m, int l, int count, int zeroBC, bool zero_start, bool exact, hipEvent_t ea,
                      hipEvent_t eb, bool faces_ready, bool first_done) {
    Level &V = m->lv[l];
    Bc bc = pressure_bc(m, l, zeroBC);
    hipStream_t s = m->stream;
    if (m->d.gsSmooth) {
        // the red-black sweeps read face ghosts only: a level whose face-ghost images are current
        // (faces_ready) needs no fill
        if (!zero_start && !faces_ready) launch_fill_ghosts<double>(V.L, V.u, bc, s);
        // a zero start without sweeps still starts from zero (poisson.cc clears the level)
        if (zero_start && count == 0) launch_fill<double>(V.u, lay_size(V.L), 0.0, s);
        if (ea) mark(m, ea);
        for (int n = 0; n < count; n++) {
            // first_done: the first (zero-start) colour-0 half sweep was done by the restriction
            if (!(first_done && n == 0))
                launch_rb_sweep<double>(V.L, V.u, V.f, V.c64[0], V.c64[1], V.c64[2], 0, bc, zero_start && n == 0, s);
            launch_rb_sweep<double>(V.L, V.u, V.f, V.c64[0], V.c64[1], V.c64[2], 1, bc, false, s);
        }
        if (eb) mark(m, eb);
    } else {
        // Jacobi: tmp(l) = 0; per sweep {imposeBC; tmp = J(u); swap hierarchies} (poisson_d3.cc:152-168)
        launch_fill<double>(V.t, lay_size(V.L), 0.0, s);
        for (int n = 0; n < count; n++) {
            Level &W = m->lv[l];
            launch_fill_ghosts<double>(W.L, W.u, bc, s);
            launch_jacobi<double>(W.L, W.u, W.f, W.t, W.c64[0], W.c64[1], W.c64[2], s);
            for (Level &X : m->lv) std::swap(X.u, X.t);
        }
    }
    if (exact || !m->d.gsSmooth) launch_fill_ghosts<double>(V.L, m->lv[l].u, bc, s);
}

// every rule of b writes images (no kept or zero ghosts)
static bool image_rules(const Bc &b) { return b.rule[0] <= G_ANTI && b.rule[1] <= G_ANTI && b.rule[2] <= G_ANTI; }

void sr::residual_level(sr_mg *m, int l, int color, bool images) {
    Level &V = m->lv[l];
    const Bc tb = tmp_bc(m);
    // with colour 0 only (read by the red-black restriction: coincident node and edge points) and
    // image rules, the residual kernel writes the face and edge images itself
    const bool img = images && color == 0 && image_rules(tb);
    launch_residual<double, double>(V.L, V.u, V.f, V.t, V.c64[0], V.c64[1], V.c64[2], m->stream, color,
                                    img ? &tb : nullptr);
    if (!img) launch_fill_ghosts<double>(V.L, V.t, tb, m->stream);
}

void sr::coarsest_level(sr_mg *m, int l, int zeroBC) {
    Level &V = m->lv[l];
    Bc bc = pressure_bc(m, l, zeroBC);
    if (m->pinv64 && l == m->depth) {
        for (int q = 0; q < 6; q++) bc.wall[q] = nullptr;
        launch_coarsest_direct<double>(V.L, V.u, V.f, m->pinv64, m->ncd, bc, m->diters, m->stream);
        return;
    }
    launch_coarsest<double>(V.L, V.u, V.f, V.c64[0], V.c64[1], V.c64[2], bc, m->d.mgTolerance, m->maxCount,
                            m->diters, m->stream);
}

// range projection of the coarsest right-hand side with that level's own null vectors
static void project_coarsest(sr_mg *m) {
    Level &Cv = m->lv[m->depth];
    launch_project_small<double>(Cv.L, Cv.f, Cv.wc[0], Cv.wc[1], Cv.wc[2], Cv.wcsum, m->stream);
}

// poisson.cc:219-280 with restriction fused into the down leg and += smoothedPres into prolong
void sr::vcycle_fp64(sr_mg *m) {
    const sr_mg_desc &d = m->d;
    int D = m->depth;
    int Lc = (m->coarse_enable && d.gsSmooth) ? m->coarse_first64 : (1 << 20);
    int bottom = Lc <= D ? Lc : D;   // last level handled by per-level launches + 1
    smooth_level(m, 0, d.preSmooth, 0, false, false, m->sw0, m->sw1);
    // pre-smoothing ends on a colour-1 half sweep, after which the colour-1 residual is zero: only
    // colour 0 is computed and restricted
    const bool red0 = d.gsSmooth && d.preSmooth > 0;
    for (int l = 0; l < bottom; l++) {
        residual_level(m, l, red0 ? 0 : -1, l > 0);
        // a level smoothed from zero by Gauss-Seidel starts with a colour-0 half sweep that reads no
        // u: the restriction that makes its right-hand side does it (ZeroStart)
        const bool fuse = d.gsSmooth && d.preSmooth > 0 && l + 1 != Lc && l + 1 != D;
        Level &Nx = m->lv[l + 1];
        const ZeroStart<double> zs{Nx.u, Nx.c64[0], Nx.c64[1], Nx.c64[2], pressure_bc(m, l + 1, 1)};
        launch_restrict<double>(m->lv[l].L, Nx.L, m->lv[l].t, Nx.f, m->stream, red0, fuse ? &zs : nullptr);
        if (l + 1 == Lc) {
            launch_coarse_cycle<double>(m->cdesc64, m->clds64, m->lv[Lc].u, m->lv[Lc].f, m->diters, m->stream);
        } else if (l + 1 == D) {
            launch_fill<double>(m->lv[D].u, lay_size(m->lv[D].L), 0.0, m->stream);
            if (d.coarseProject && singular(m)) project_coarsest(m);
            coarsest_level(m, D, 1);
        } else if (d.gsSmooth) {
            smooth_level(m, l + 1, d.preSmooth, 1, true, false, nullptr, nullptr, false, fuse);
        } else {
            launch_fill<double>(m->lv[l + 1].u, lay_size(m->lv[l + 1].L), 0.0, m->stream);
            smooth_level(m, l + 1, d.preSmooth, 1, false, false);
        }
    }
    // post-smoothing sweeps colour 0 first, which reads only colour-1 nodes: the correction of the
    // colour-0 nodes would be overwritten unread (Gauss-Seidel with at least one sweep)
    const int pcol = d.gsSmooth && d.postSmooth > 0 ? 1 : -1;
    // below level 0 the prolongation writes the face-ghost images of what it updates (every ghost
    // rule there is an image rule), so the post-smoothing starts without a ghost fill; level 0 keeps
    // the separate fill (measured faster there)
    for (int l = bottom - 1; l >= 0; l--) {
        const bool img = l > 0 && d.gsSmooth;
        const Bc bc = pressure_bc(m, l, 1);
        launch_prolong<double, double>(m->lv[l].L, m->lv[l + 1].L, m->lv[l + 1].u, m->lv[l].u, 1, m->stream, pcol,
                                       img ? &bc : nullptr);
        smooth_level(m, l, d.postSmooth, l == 0 ? 0 : 1, false, false, nullptr, nullptr, img);
    }
}

// ------------------------------------------------------------------------------------------------
// FP32 hierarchy (mixed precision): the V-cycle on the error equation A e = r with e = 0 at start.
// Level-0 ghosts follow the BCs of the error: periodic / mirror as for u, Dirichlet walls -> 0.
// ------------------------------------------------------------------------------------------------
Bc sr::error_bc(const sr_mg *m, int l, int zeroBC) {
    Bc b = pressure_bc(m, l, zeroBC);
    for (int q = 0; q < 6; q++) b.wall[q] = nullptr;
    return b;
}

static void smooth32(sr_mg *m, int l, int count, int zeroBC, bool zero_start, hipEvent_t ea = nullptr,
                     hipEvent_t eb = nullptr, bool faces_ready = false, bool first_done = false) {
    Level &V = m->lv[l];
    Bc bc = error_bc(m, l, zeroBC);
    bool wall = bc.rule[0] == G_WALL || bc.rule[1] == G_WALL || bc.rule[2] == G_WALL;
    bool zs = zero_start && !wall;
    hipStream_t s = m->stream;
    if (!zs && !(faces_ready && !zero_start)) {
        if (zero_start) launch_fill<float>(V.u32, lay_size(V.L), 0.0f, s);
        launch_fill_ghosts<float>(V.L, V.u32, bc, s);
    } else if (count == 0) {
        launch_fill<float>(V.u32, lay_size(V.L), 0.0f, s);   // a zero start without sweeps
    }
    if (ea) mark(m, ea);
    for (int n = 0; n < count; n++) {
        // first_done: the first (zero-start) colour-0 half sweep was done by the restriction
        if (!(first_done && n == 0))
            launch_rb_sweep<float>(V.L, V.u32, V.f32, V.c32[0], V.c32[1], V.c32[2], 0, bc, zs && n == 0, s);
        launch_rb_sweep<float>(V.L, V.u32, V.f32, V.c32[0], V.c32[1], V.c32[2], 1, bc, false, s);
    }
    if (eb) mark(m, eb);
}

void sr::vcycle_fp32(sr_mg *m) {
    const sr_mg_desc &d = m->d;
    int D = m->depth;
    hipStream_t s = m->stream;
    int Lc = m->coarse_enable ? m->coarse_first32 : (1 << 20);
    int bottom = Lc <= D ? Lc : D;
    smooth32(m, 0, m->pre32, 0, true, m->sw0, m->sw1);
    // the pre-smoothing ends on colour 1: the colour-1 residual is zero, only colour 0 is computed
    const bool red0 = m->pre32 > 0;
    for (int l = 0; l < bottom; l++) {
        Level &V = m->lv[l];
        // below level 0 the colour-0 residual writes its face and edge images (image rules only)
        const Bc tb = tmp_bc(m);
        const bool img = l > 0 && red0 && image_rules(tb);
        launch_residual<float, float>(V.L, V.u32, V.f32, V.t32, V.c32[0], V.c32[1], V.c32[2], s, red0 ? 0 : -1,
                                      img ? &tb : nullptr);
        if (!img) launch_fill_ghosts<float>(V.L, V.t32, tb, s);
        // the next level's zero-start colour-0 half sweep done by its restriction (as in the FP64 cycle)
        Level &Nx = m->lv[l + 1];
        const Bc nb = error_bc(m, l + 1, 1);
        const bool nwall = nb.rule[0] == G_WALL || nb.rule[1] == G_WALL || nb.rule[2] == G_WALL;
        const bool fuse = m->pre32 > 0 && !nwall && l + 1 != Lc && l + 1 != D;
        const ZeroStart<float> zs{Nx.u32, Nx.c32[0], Nx.c32[1], Nx.c32[2], nb};
        launch_restrict<float>(V.L, Nx.L, V.t32, Nx.f32, s, red0, fuse ? &zs : nullptr);
        if (l + 1 == Lc) {
            launch_coarse_cycle<float>(m->cdesc32, m->clds32, m->lv[Lc].u32, m->lv[Lc].f32, m->diters, s);
        } else if (l + 1 == D) {
            Level &Cv = m->lv[D];
            if (m->pinv32) {
                launch_coarsest_direct<float>(Cv.L, Cv.u32, Cv.f32, m->pinv32, m->ncd, error_bc(m, D, 1), m->diters, s);
                continue;
            }
            launch_fill<float>(Cv.u32, lay_size(Cv.L), 0.0f, s);
            if (d.coarseProject && singular(m))
                launch_project_small<float>(Cv.L, Cv.f32, Cv.wc[0], Cv.wc[1], Cv.wc[2], Cv.wcsum, s);
            launch_coarsest<float>(Cv.L, Cv.u32, Cv.f32, Cv.c32[0], Cv.c32[1], Cv.c32[2], error_bc(m, D, 1),
                                   d.mgTolerance, m->maxCount, m->diters, s);
        } else {
            smooth32(m, l + 1, m->pre32, 1, true, nullptr, nullptr, false, fuse);
        }
    }
    // colour 0 is swept first by the post-smoothing: its correction would be overwritten unread
    const int pcol = m->post32 > 0 ? 1 : -1;
    // below level 0 the prolongation writes the face-ghost images of what it updates, so the
    // post-smoothing starts without a ghost fill (as in the FP64 cycle)
    for (int l = bottom - 1; l >= 0; l--) {
        const Bc bc = error_bc(m, l, 1);
        launch_prolong<float, float>(m->lv[l].L, m->lv[l + 1].L, m->lv[l + 1].u32, m->lv[l].u32, 1, s, pcol,
                                     l > 0 ? &bc : nullptr);
        smooth32(m, l, m->post32, l == 0 ? 0 : 1, false, nullptr, nullptr, l > 0);
    }
}

int sr::ensure_fp32(sr_mg *m) {
    for (Level &V : m->lv) {
        if (V.u32) continue;
        long sz = lay_size(V.L);
        HIPCHK(hipMalloc((void **)&V.u32, sizeof(float) * sz));
        HIPCHK(hipMalloc((void **)&V.f32, sizeof(float) * sz));
        HIPCHK(hipMalloc((void **)&V.t32, sizeof(float) * sz));
        launch_fill<float>(V.u32, sz, 0.0f, m->stream);
        launch_fill<float>(V.f32, sz, 0.0f, m->stream);
        launch_fill<float>(V.t32, sz, 0.0f, m->stream);
    }
    if (coarsest_lds_bytes<float>(m->lv[m->depth].L) > 60 * 1024) return set_err(SR_ERR_ARG, "coarsest level too large");
    return SR_OK;
}

void sr::norm_level0(sr_mg *m, int color) {
    Level &V = m->lv[0];
    launch_norm_res<double>(V.L, V.u, V.f, V.c64[0], V.c64[1], V.c64[2], m->part2, m->fterms, m->dres, m->stream,
                            color);
}

void sr::f_terms_level0(sr_mg *m) {
    Level &V = m->lv[0];
    launch_f_terms<double>(V.L, V.f, m->part, m->fterms, m->stream);
}

// computeError from the four reduced quantities (poisson_d3.cc:433-458)
double sr::error_from(const sr_mg *m, const double *q, int norm) {
    const Lay &L = m->lv[0].L;
    double cnt = (double)L.nx * L.ny * L.nz;
    if (norm == 1) return q[3] != 0.0 ? q[2] / q[3] : q[2];
    if (norm == 2) return q[1] != 0.0 ? std::sqrt(q[0] / cnt) / std::sqrt(q[1] / cnt) : std::sqrt(q[0] / cnt);
    return 0.0;
}

// ------------------------------------------------------------------------------------------------
// per-operator entry points
// ------------------------------------------------------------------------------------------------
static double *level_array(sr_mg *m, int which, int level) {
    Level &V = m->lv[level];
    return which == SR_PRESSURE ? V.u : which == SR_RESIDUAL ? V.f : V.t;
}

extern "C" int sr_mg_set_level(sr_mg *m, int level) {
    if (!m || level < 0 || level > m->depth) return set_err(SR_ERR_ARG, "level out of range");
    m->vLevel = level;
    return SR_OK;
}
extern "C" int sr_mg_get_level(const sr_mg *m, int *level) {
    if (!m || !level) return set_err(SR_ERR_ARG, "null argument");
    *level = m->vLevel;
    return SR_OK;
}
extern "C" int sr_mg_set_zero_bc(sr_mg *m, int z) {
    if (!m) return set_err(SR_ERR_ARG, "null argument");
    m->zeroBC = z ? 1 : 0;
    return SR_OK;
}
extern "C" int sr_mg_level_dims(const sr_mg *m, int level, int *nx, int *ny, int *nz) {
    if (!m || level < 0 || level > m->depth) return set_err(SR_ERR_ARG, "level out of range");
    *nx = m->lv[level].L.nx;
    *ny = m->lv[level].L.ny;
    *nz = m->lv[level].L.nz;
    return SR_OK;
}

static int ensure_stage(sr_mg *m) {
    if (m->stage) return SR_OK;
    const Lay &L = m->lv[0].L;
    size_t n = (size_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    HIPCHK(hipMalloc((void **)&m->stage, sizeof(double) * n));
    HIPCHK(hipMalloc((void **)&m->stage2, sizeof(double) * n));
    return SR_OK;
}

extern "C" int sr_mg_upload(sr_mg *m, int which, int level, const double *host) {
    if (!m || !host || level < 0 || level > m->depth || which < 0 || which > 2)
        return set_err(SR_ERR_ARG, "bad upload arguments");
    RETCHK(ensure_stage(m));
    const Lay &L = m->lv[level].L;
    size_t n = (size_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    HIPCHK(hipMemcpyAsync(m->stage, host, sizeof(double) * n, hipMemcpyHostToDevice, m->stream));
    launch_pack<double, double>(L, m->stage, level_array(m, which, level), 0, m->stream);
    HIPCHK(hipStreamSynchronize(m->stream));
    return SR_OK;
}

extern "C" int sr_mg_download(sr_mg *m, int which, int level, double *host) {
    if (!m || !host || level < 0 || level > m->depth || which < 0 || which > 2)
        return set_err(SR_ERR_ARG, "bad download arguments");
    RETCHK(ensure_stage(m));
    const Lay &L = m->lv[level].L;
    size_t n = (size_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    launch_unpack<double, double>(L, level_array(m, which, level), m->stage, m->stream);
    HIPCHK(hipMemcpyAsync(host, m->stage, sizeof(double) * n, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    return SR_OK;
}

static int sync_check(sr_mg *m) {
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(m->stream));
    return SR_OK;
}

static int wall_guard(const sr_mg *m) {
    if (m->d.testPoisson && !m->zeroBC && m->vLevel != 0)
        return set_err(SR_ERR_STATE, "non-zero Dirichlet walls exist on level 0 only");
    return SR_OK;
}

extern "C" int sr_mg_op_smooth(sr_mg *m, int count) {
    if (!m || count < 0) return set_err(SR_ERR_ARG, "bad smooth arguments");
    RETCHK(wall_guard(m));
    smooth_level(m, m->vLevel, count, m->zeroBC, false, true);
    return sync_check(m);
}

extern "C" int sr_mg_op_compute_residual(sr_mg *m) {
    if (!m) return set_err(SR_ERR_ARG, "null");
    residual_level(m, m->vLevel);
    return sync_check(m);
}

extern "C" int sr_mg_op_coarsen(sr_mg *m) {
    if (!m || m->vLevel >= m->depth) return set_err(SR_ERR_STATE, "coarsen below the coarsest level");
    int l = m->vLevel;
    launch_restrict<double>(m->lv[l].L, m->lv[l + 1].L, m->lv[l].t, m->lv[l + 1].f, m->stream);
    m->vLevel += 1;
    return sync_check(m);
}

extern "C" int sr_mg_op_prolong(sr_mg *m) {
    if (!m || m->vLevel < 1) return set_err(SR_ERR_STATE, "prolong above the finest level");
    int l = m->vLevel - 1;
    // pressureData(vLevel) = 0.0 over the padded level, then the core (poisson_d3.cc:307-309)
    launch_fill<double>(m->lv[l].u, lay_size(m->lv[l].L), 0.0, m->stream);
    launch_prolong<double, double>(m->lv[l].L, m->lv[l + 1].L, m->lv[l + 1].u, m->lv[l].u, 0, m->stream);
    m->vLevel = l;
    return sync_check(m);
}

extern "C" int sr_mg_op_solve_coarsest(sr_mg *m) {
    if (!m) return set_err(SR_ERR_ARG, "null");
    if (coarsest_lds_bytes<double>(m->lv[m->vLevel].L) > 60 * 1024)
        return set_err(SR_ERR_ARG, "level too large for the single-workgroup solve");
    RETCHK(wall_guard(m));
    coarsest_level(m, m->vLevel, m->zeroBC);
    return sync_check(m);
}

extern "C" int sr_mg_op_impose_bc(sr_mg *m) {
    if (!m) return set_err(SR_ERR_ARG, "null");
    RETCHK(wall_guard(m));
    Level &V = m->lv[m->vLevel];
    launch_fill_ghosts<double>(V.L, V.u, pressure_bc(m, m->vLevel, m->zeroBC), m->stream);
    return sync_check(m);
}

extern "C" int sr_mg_op_update_pads(sr_mg *m, int which) {
    if (!m || which < 0 || which > 2) return set_err(SR_ERR_ARG, "bad array");
    Level &V = m->lv[m->vLevel];
    launch_fill_ghosts<double>(V.L, level_array(m, which, m->vLevel), pads_bc(m), m->stream);
    return sync_check(m);
}

extern "C" int sr_mg_op_compute_error(sr_mg *m, int norm, double *out) {
    if (!m || !out) return set_err(SR_ERR_ARG, "null");
    if (m->vLevel != 0) return set_err(SR_ERR_STATE, "computeError runs on the finest level");
    f_terms_level0(m);
    norm_level0(m);
    HIPCHK(hipMemcpyAsync(m->hres, m->dres, sizeof(double) * 4, hipMemcpyDeviceToHost, m->stream));
    RETCHK(sync_check(m));
    *out = error_from(m, m->hres, norm);
    return SR_OK;
}

extern "C" int sr_mg_op_vcycle(sr_mg *m) {
    if (!m) return set_err(SR_ERR_ARG, "null");
    vcycle_fp64(m);
    m->vLevel = 0;
    m->zeroBC = 0;
    return sync_check(m);
}

extern "C" int sr_mg_info(const sr_mg *m, int what, long *value) {
    if (!m || !value) return set_err(SR_ERR_ARG, "null argument");
    switch (what) {
        case 0: *value = m->coarse_first64; break;
        case 1: *value = m->coarse_first32; break;
        case 2: *value = (long)m->clds64; break;
        case 3: *value = (long)m->clds32; break;
        case 4: *value = m->lds_limit; break;
        case 5: {
            int it = 0;
            HIPCHK(hipMemcpy(&it, m->diters, sizeof(int), hipMemcpyDeviceToHost));
            *value = it;
            break;
        }
        default: return set_err(SR_ERR_ARG, "unknown info key");
    }
    return SR_OK;
}

extern "C" int sr_mg_history(const sr_mg *m, double *out, int cap) {
    if (!m) return -1;
    int n = std::min<int>(cap, (int)m->history.size());
    for (int i = 0; i < n; i++) out[i] = m->history[i];
    return n;
}

// ------------------------------------------------------------------------------------------------
// device memory helpers
// ------------------------------------------------------------------------------------------------
extern "C" int sr_device_alloc(unsigned long bytes, void **ptr) {
    if (!ptr || bytes == 0) return set_err(SR_ERR_ARG, "bad allocation request");
    HIPCHK(hipMalloc(ptr, bytes));
    return SR_OK;
}

extern "C" int sr_device_free(void *ptr) {
    if (ptr) HIPCHK(hipFree(ptr));
    return SR_OK;
}

extern "C" int sr_memcpy(void *dst, const void *src, unsigned long bytes, int kind) {
    hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost
                    : kind == 3 ? hipMemcpyDeviceToDevice : hipMemcpyDefault;
    if (kind < 1 || kind > 3 || !dst || !src) return set_err(SR_ERR_ARG, "bad copy request");
    HIPCHK(hipMemcpy(dst, src, bytes, k));
    return SR_OK;
}
