// C ABI of the hydro time step (include/saras_ts.h): hydro_d3::timeAdvance
// (src/solvers/hydro/hydro_d3.cc:254-368) with every field resident in HBM and the multigrid
// pressure solve of saras_mg.h inside.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/saras_ts.h"
#include "mg_state.h"
#include "ts_kernels.h"
#include "ghost_waves.h"

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

struct sr_ts {
    sr_ts_desc d{};
    int N[3] = {0, 0, 0};
    double h[3] = {0, 0, 0};
    FShape fp{}, fv[3]{};
    std::vector<double> hmet[3][2][4];   // host metrics, [dim][node 0 / face 1][quantity]
    double *dmet = nullptr;
    Metrics M{};
    double *V[3] = {nullptr, nullptr, nullptr};
    double *H[3] = {nullptr, nullptr, nullptr};
    double *G[3] = {nullptr, nullptr, nullptr};   // Jacobi target, swapped with V after each sweep
    double *P = nullptr, *Pp = nullptr, *mgRHS = nullptr, *divV = nullptr;
    double *scratch = nullptr;                    // derivative output of the parity entry point
    double *part = nullptr, *dred = nullptr, *hred = nullptr;
    unsigned long long *dmax = nullptr, *hmax = nullptr;
    hipEvent_t jev = nullptr;                    // marks a Jacobi residual's read-back
    std::vector<GhostWave> vbc[3];               // imposeU/V/WBCs as ordered waves
    GhostWave sync_v[3]{}, sync_p{};             // syncData of the velocity components and of P
    sr_mg *mg = nullptr;
    hipStream_t s = nullptr;
    int jac[4] = {0, 0, 0, 0};
    int mg_cycles = 0;
    // scalar_d3
    double *T = nullptr, *Trhs = nullptr, *Tg = nullptr;
    std::vector<GhostWave> tbc;
    double nu = 0.0, kappa = 0.0, Fb = 0.0;
    // learned initial guess of the pressure solve
    sr_guess *guess = nullptr;
    int train = 0;
    double *Pp1 = nullptr, *Pp2 = nullptr;   // previous two corrections
    long nsolved = 0;
};

static long fsize(const FShape &f) { return (long)(f.n[0] + 2) * (f.n[1] + 2) * (f.n[2] + 2); }

// ------------------------------------------------------------------------------------------------
// ghost waves
// ------------------------------------------------------------------------------------------------
// hydro::imposeU/V/WBCs (hydro.cc:386-461) with the walls of hydro::initVBC (hydro.cc:287-371):
// syncData, then x walls, y walls, and the top and bottom walls.  Walls normal to the component are
// no-penetration (dirichletFC 0), the others no-slip (dirichletCC 0; the lid of probType 1 moves
// Vx at 1); periodicCC / periodicFC in z when zPer.
static std::vector<GhostWave> vbc_waves(const sr_ts_desc &d, const FShape &f, int c) {
    std::vector<GhostWave> ws;
    ws.push_back(sync_wave(f, d.xPer, d.yPer));
    int per[3] = {d.xPer, d.yPer, d.zPer};
    for (int dim = 0; dim < 2; dim++) {
        if (per[dim]) continue;
        int kind = dim == c ? 3 : 2;
        GhostWave w{};
        add_op(w, wall_op(f, 0, 2 * dim, kind, 0.0));
        add_op(w, wall_op(f, 0, 2 * dim + 1, kind, 0.0));
        ws.push_back(w);
    }
    GhostWave w{};
    if (per[2]) {
        int kind = c == 2 ? 1 : 0;
        add_op(w, wall_op(f, 1, 5, kind, 0.0));
        add_op(w, wall_op(f, 1, 4, kind, 0.0));
    } else {
        int kind = c == 2 ? 3 : 2;
        double top = (d.probType == 1 && c == 0) ? 1.0 : 0.0;
        add_op(w, wall_op(f, 0, 5, kind, top));
        add_op(w, wall_op(f, 0, 4, kind, 0.0));
    }
    ws.push_back(w);
    return ws;
}

static void impose_vbc(sr_ts *t, int c) {
    for (const GhostWave &w : t->vbc[c]) ts_ghost_wave(t->fv[c], t->V[c], w, t->s);
}

// scalar::imposeTBCs (scalar.cc:240-255) with the walls of initTBC (scalar.cc:196-228): syncData,
// adiabatic (neumannCC) x and y walls, then top and bottom plates (probType 5: cold top, hot
// bottom; 6: the reverse), or periodicCC in z
static std::vector<GhostWave> tbc_waves(const sr_ts_desc &d, const FShape &f) {
    std::vector<GhostWave> ws;
    ws.push_back(sync_wave(f, d.xPer, d.yPer));
    int per[3] = {d.xPer, d.yPer, d.zPer};
    for (int dim = 0; dim < 2; dim++) {
        if (per[dim]) continue;
        GhostWave w{};
        add_op(w, wall_op(f, 0, 2 * dim, 4, 0.0));
        add_op(w, wall_op(f, 0, 2 * dim + 1, 4, 0.0));
        ws.push_back(w);
    }
    GhostWave w{};
    if (per[2]) {
        add_op(w, wall_op(f, 1, 5, 0, 0.0));
        add_op(w, wall_op(f, 1, 4, 0, 0.0));
    } else {
        double top = d.probType == 6 ? 1.0 : 0.0, bottom = d.probType == 6 ? 0.0 : 1.0;
        add_op(w, wall_op(f, 0, 5, 2, top));
        add_op(w, wall_op(f, 0, 4, 2, bottom));
    }
    ws.push_back(w);
    return ws;
}

static void impose_tbc(sr_ts *t) {
    for (const GhostWave &w : t->tbc) ts_ghost_wave(t->fp, t->T, w, t->s);
}

// ------------------------------------------------------------------------------------------------
// construction
// ------------------------------------------------------------------------------------------------
static int create_impl(const sr_ts_desc *desc, sr_ts *t) {
    t->d = *desc;
    const sr_ts_desc &d = t->d;
    int ind[3] = {d.xInd, d.yInd, d.zInd};
    double L[3] = {d.Lx, d.Ly, d.Lz};
    int gt[3] = {d.xGrid, d.yGrid, d.zGrid};
    double beta[3] = {d.betaX, d.betaY, d.betaZ};
    for (int k = 0; k < 3; k++) {
        if (ind[k] < 2 || ind[k] > 12) return set_err(SR_ERR_ARG, "grid size index out of range 2..12");
        if (L[k] <= 0.0) return set_err(SR_ERR_ARG, "domain length must be positive");
    }
    if (d.dt <= 0.0 || d.cnTolerance <= 0.0) return set_err(SR_ERR_ARG, "dt and cnTolerance must be positive");
    if (!d.scalar && d.Re <= 0.0) return set_err(SR_ERR_ARG, "Re must be positive");
    if (d.scalar) {
        if (d.Ra <= 0.0 || d.Pr <= 0.0) return set_err(SR_ERR_ARG, "Ra and Pr must be positive");
        // scalar.cc:69-86, buoyantForce.cc:45-55
        switch (d.rbcType) {
            case 1: t->nu = d.Pr; t->kappa = 1.0; t->Fb = d.Ra * d.Pr; break;
            case 2: t->nu = std::sqrt(d.Pr / d.Ra); t->kappa = 1.0 / std::sqrt(d.Pr * d.Ra); t->Fb = 1.0; break;
            case 3: t->nu = 1.0; t->kappa = 1.0 / d.Pr; t->Fb = d.Ra; break;
            case 4: t->nu = std::sqrt(d.Pr / d.Ra); t->kappa = 1.0 / std::sqrt(d.Pr * d.Ra); t->Fb = d.Pr; break;
            default: return set_err(SR_ERR_ARG, "rbcType must be 1..4");
        }
    } else {
        t->nu = 1.0 / d.Re;
    }
    for (int k = 0; k < 3; k++) {
        int n = (1 << ind[k]) + 1;
        t->N[k] = n;
        t->h[k] = 1.0 / (double)(n - 1);
        for (int face = 0; face < 2; face++)
            for (int q = 0; q < 4; q++) t->hmet[k][face][q].assign((size_t)(face ? n + 1 : n + 2), 0.0);
        for (int face = 0; face < 2; face++) {
            int r = sr_grid_axis(n, L[k], gt[k], beta[k], face, t->hmet[k][face][0].data(), t->hmet[k][face][1].data(),
                                 t->hmet[k][face][2].data(), t->hmet[k][face][3].data());
            if (r != SR_OK) return set_err(r, "invalid grid description");
        }
    }
    for (int k = 0; k < 3; k++) {
        t->fp.n[k] = t->N[k];
        t->fp.s[k] = 1;
    }
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < 3; k++) {
            t->fv[c].n[k] = c == k ? t->N[k] - 1 : t->N[k];
            t->fv[c].s[k] = c == k ? 0 : 1;
        }
    // metric arrays in one allocation
    size_t tot = 0;
    for (int k = 0; k < 3; k++)
        for (int face = 0; face < 2; face++)
            for (int q = 0; q < 4; q++) tot += t->hmet[k][face][q].size();
    HIPCHK(hipMalloc((void **)&t->dmet, sizeof(double) * tot));
    size_t off = 0;
    for (int k = 0; k < 3; k++)
        for (int face = 0; face < 2; face++)
            for (int q = 0; q < 4; q++) {
                const std::vector<double> &v = t->hmet[k][face][q];
                HIPCHK(hipMemcpy(t->dmet + off, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
                t->M.m[k][face][q] = t->dmet + off;
                off += v.size();
            }
    // the pressure solver on the node grid (hydro_d3.cc:63-66)
    sr_mg_desc md;
    std::memset(&md, 0, sizeof(md));
    md.xInd = d.xInd; md.yInd = d.yInd; md.zInd = d.zInd;
    md.vcDepth = d.vcDepth; md.vcCount = d.vcCount;
    md.preSmooth = d.preSmooth; md.postSmooth = d.postSmooth;
    md.gsSmooth = d.gsSmooth; md.resType = d.resType;
    md.xPer = d.xPer; md.yPer = d.yPer; md.zPer = d.zPer;
    md.semantics = d.semantics;
    md.precision = d.precision;
    md.coarseDirect = d.coarseDirect;
    md.innerPreSmooth = d.innerPreSmooth;
    md.innerPostSmooth = d.innerPostSmooth;
    md.padMode = d.padMode;
    md.mgTolerance = d.mgTolerance;
    md.relTolerance = d.relTolerance > 0.0 ? d.relTolerance : 1e-10;
    md.dXi = t->h[0]; md.dEt = t->h[1]; md.dZt = t->h[2];
    md.xix2 = t->hmet[0][0][3].data(); md.xixx = t->hmet[0][0][2].data();
    md.ety2 = t->hmet[1][0][3].data(); md.etyy = t->hmet[1][0][2].data();
    md.ztz2 = t->hmet[2][0][3].data(); md.ztzz = t->hmet[2][0][2].data();
    RETCHK(sr_mg_create(&md, &t->mg));
    t->s = t->mg->stream;
    // fields
    long np = fsize(t->fp);
    HIPCHK(hipMalloc((void **)&t->P, sizeof(double) * np));
    HIPCHK(hipMalloc((void **)&t->Pp, sizeof(double) * np));
    HIPCHK(hipMalloc((void **)&t->mgRHS, sizeof(double) * np));
    HIPCHK(hipMalloc((void **)&t->divV, sizeof(double) * np));
    long nmax = np;
    for (int c = 0; c < 3; c++) {
        long nv = fsize(t->fv[c]);
        nmax = std::max(nmax, nv);
        HIPCHK(hipMalloc((void **)&t->V[c], sizeof(double) * nv));
        HIPCHK(hipMalloc((void **)&t->H[c], sizeof(double) * nv));
        HIPCHK(hipMalloc((void **)&t->G[c], sizeof(double) * nv));
        ts_fill(t->V[c], nv, 0.0, t->s);
        ts_fill(t->H[c], nv, 0.0, t->s);
        ts_fill(t->G[c], nv, 0.0, t->s);
    }
    HIPCHK(hipMalloc((void **)&t->scratch, sizeof(double) * nmax));
    ts_fill(t->P, np, 0.0, t->s);
    ts_fill(t->Pp, np, 0.0, t->s);
    ts_fill(t->mgRHS, np, 0.0, t->s);
    ts_fill(t->divV, np, 0.0, t->s);
    if (d.scalar) {
        HIPCHK(hipMalloc((void **)&t->T, sizeof(double) * np));
        HIPCHK(hipMalloc((void **)&t->Trhs, sizeof(double) * np));
        HIPCHK(hipMalloc((void **)&t->Tg, sizeof(double) * np));
        ts_fill(t->T, np, 0.0, t->s);
        ts_fill(t->Trhs, np, 0.0, t->s);
        ts_fill(t->Tg, np, 0.0, t->s);
        t->tbc = tbc_waves(d, t->fp);
    }
    size_t nparts = std::max<size_t>(2 * (size_t)ts_energy_parts(t->fp), ts_jacobi_res_parts());
    nparts = std::max<size_t>(nparts, (size_t)ts_jacobi_vres_parts(t->fp));
    for (int c = 0; c < 3; c++) nparts = std::max<size_t>(nparts, (size_t)ts_jacobi_vres_parts(t->fv[c]));
    HIPCHK(hipMalloc((void **)&t->part, sizeof(double) * nparts));
    HIPCHK(hipEventCreateWithFlags(&t->jev, hipEventDisableTiming));
    HIPCHK(hipMalloc((void **)&t->dred, sizeof(double) * 4));
    HIPCHK(hipHostMalloc((void **)&t->hred, sizeof(double) * 4, hipHostMallocDefault));
    HIPCHK(hipMalloc((void **)&t->dmax, sizeof(unsigned long long)));
    HIPCHK(hipHostMalloc((void **)&t->hmax, sizeof(unsigned long long), hipHostMallocDefault));
    for (int c = 0; c < 3; c++) {
        t->vbc[c] = vbc_waves(d, t->fv[c], c);
        t->sync_v[c] = sync_wave(t->fv[c], d.xPer, d.yPer);
    }
    t->sync_p = sync_wave(t->fp, d.xPer, d.yPer);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(t->s));
    return SR_OK;
}

extern "C" int sr_ts_destroy(sr_ts *t) {
    if (!t) return SR_OK;
    if (t->s) (void)hipStreamSynchronize(t->s);
    double *bufs[] = {t->dmet, t->P,    t->Pp,   t->mgRHS, t->divV, t->scratch, t->part,
                      t->dred, t->T,    t->Trhs, t->Tg,    t->Pp1,  t->Pp2};
    for (double *b : bufs)
        if (b) (void)hipFree(b);
    for (int c = 0; c < 3; c++) {
        if (t->V[c]) (void)hipFree(t->V[c]);
        if (t->H[c]) (void)hipFree(t->H[c]);
        if (t->G[c]) (void)hipFree(t->G[c]);
    }
    if (t->dmax) (void)hipFree(t->dmax);
    if (t->jev) (void)hipEventDestroy(t->jev);
    if (t->hred) (void)hipHostFree(t->hred);
    if (t->hmax) (void)hipHostFree(t->hmax);
    if (t->mg) sr_mg_destroy(t->mg);
    delete t;
    return SR_OK;
}

extern "C" int sr_ts_create(const sr_ts_desc *desc, sr_ts **out) {
    if (!desc || !out) return set_err(SR_ERR_ARG, "null argument");
    *out = nullptr;
    sr_ts *t = new sr_ts();
    int r = create_impl(desc, t);
    if (r != SR_OK) {
        std::string msg = sr_last_error();
        sr_ts_destroy(t);
        return set_err(r, msg);
    }
    *out = t;
    return SR_OK;
}

// ------------------------------------------------------------------------------------------------
// field access
// ------------------------------------------------------------------------------------------------
static int field(sr_ts *t, int which, double **ptr, const FShape **shape) {
    switch (which) {
        case SR_TS_VX: case SR_TS_VY: case SR_TS_VZ:
            *ptr = t->V[which]; *shape = &t->fv[which]; return SR_OK;
        case SR_TS_P: *ptr = t->P; *shape = &t->fp; return SR_OK;
        case SR_TS_PP: *ptr = t->Pp; *shape = &t->fp; return SR_OK;
        case SR_TS_MGRHS: *ptr = t->mgRHS; *shape = &t->fp; return SR_OK;
        case SR_TS_HX: case SR_TS_HY: case SR_TS_HZ:
            *ptr = t->H[which - SR_TS_HX]; *shape = &t->fv[which - SR_TS_HX]; return SR_OK;
        case SR_TS_DIV: *ptr = t->divV; *shape = &t->fp; return SR_OK;
        case SR_TS_PP1:
        case SR_TS_PP2:
            if (!t->Pp1) return set_err(SR_ERR_STATE, "no correction history (no guess attached)");
            *ptr = which == SR_TS_PP1 ? t->Pp1 : t->Pp2;
            *shape = &t->fp;
            return SR_OK;
        case SR_TS_T:
        case SR_TS_TRHS:
            if (!t->T) return set_err(SR_ERR_STATE, "no temperature field (desc.scalar = 0)");
            *ptr = which == SR_TS_T ? t->T : t->Trhs;
            *shape = &t->fp;
            return SR_OK;
        default: return set_err(SR_ERR_ARG, "unknown field");
    }
}

extern "C" int sr_ts_field_dims(const sr_ts *t, int which, int *nx, int *ny, int *nz) {
    if (!t || !nx || !ny || !nz) return set_err(SR_ERR_ARG, "null argument");
    double *p;
    const FShape *f;
    RETCHK(field(const_cast<sr_ts *>(t), which, &p, &f));
    *nx = f->n[0]; *ny = f->n[1]; *nz = f->n[2];
    return SR_OK;
}

extern "C" int sr_ts_field_ptr(sr_ts *t, int which, double **dev) {
    if (!t || !dev) return set_err(SR_ERR_ARG, "null argument");
    const FShape *f;
    return field(t, which, dev, &f);
}

extern "C" int sr_ts_upload(sr_ts *t, int which, const double *host) {
    if (!t || !host) return set_err(SR_ERR_ARG, "null argument");
    double *p;
    const FShape *f;
    RETCHK(field(t, which, &p, &f));
    HIPCHK(hipMemcpyAsync(p, host, sizeof(double) * fsize(*f), hipMemcpyHostToDevice, t->s));
    HIPCHK(hipStreamSynchronize(t->s));
    return SR_OK;
}

extern "C" int sr_ts_download(sr_ts *t, int which, double *host) {
    if (!t || !host) return set_err(SR_ERR_ARG, "null argument");
    double *p;
    const FShape *f;
    RETCHK(field(t, which, &p, &f));
    HIPCHK(hipMemcpyAsync(host, p, sizeof(double) * fsize(*f), hipMemcpyDeviceToHost, t->s));
    HIPCHK(hipStreamSynchronize(t->s));
    return SR_OK;
}

// ------------------------------------------------------------------------------------------------
// pieces of the step
// ------------------------------------------------------------------------------------------------
extern "C" int sr_ts_init_tgv(sr_ts *t) {
    if (!t) return set_err(SR_ERR_ARG, "null argument");
    ts_fill(t->P, fsize(t->fp), 1.0, t->s);
    ts_fill(t->Pp, fsize(t->fp), 0.0, t->s);
    for (int c = 0; c < 3; c++) ts_tgv(t->fv[c], t->M, c, t->d.Lx, t->d.Ly, t->d.Lz, t->V[c], t->s);
    for (int c = 0; c < 3; c++) impose_vbc(t, c);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(t->s));
    return SR_OK;
}

extern "C" int sr_ts_impose_vbc(sr_ts *t, int component) {
    if (!t || component < 0 || component > 2) return set_err(SR_ERR_ARG, "bad component");
    impose_vbc(t, component);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(t->s));
    return SR_OK;
}

extern "C" int sr_ts_deriv(sr_ts *t, int which, int order, int dim, double *out) {
    if (!t || !out || (order != 1 && order != 2) || dim < 0 || dim > 2) return set_err(SR_ERR_ARG, "bad argument");
    if (which != SR_TS_VX && which != SR_TS_VY && which != SR_TS_VZ && which != SR_TS_P)
        return set_err(SR_ERR_ARG, "derivatives are taken of a velocity component or P");
    double *p;
    const FShape *f;
    RETCHK(field(t, which, &p, &f));
    ts_fill(t->scratch, fsize(*f), 0.0, t->s);
    ts_deriv(*f, t->M, p, t->scratch, order, dim, t->h[dim], t->s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, t->scratch, sizeof(double) * fsize(*f), hipMemcpyDeviceToHost, t->s));
    HIPCHK(hipStreamSynchronize(t->s));
    return SR_OK;
}

static void build_rhs(sr_ts *t) {
    const double *Vc[3] = {t->V[0], t->V[1], t->V[2]};
    // hydro_d3 multiplies the diffusion by 1/Re (hydro_d3.cc:259-262), scalar_d3 by nu
    const double visc = t->d.scalar ? t->nu : 1.0 / t->d.Re;
    for (int c = 0; c < 3; c++) {
        ts_build_rhs(c, t->fv, t->fp, t->M, Vc, t->P, t->H[c], t->h, visc, t->d.dt, t->d.scalar ? t->T : nullptr,
                     t->Fb, t->s);
        ts_ghost_wave(t->fv[c], t->H[c], t->sync_v[c], t->s);
    }
}

extern "C" int sr_ts_build_rhs(sr_ts *t) {
    if (!t) return set_err(SR_ERR_ARG, "null argument");
    build_rhs(t);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(t->s));
    return SR_OK;
}

static JacobiCoef jacobi_coef(const sr_ts *t) { return jacobi_coef(t->h, t->d.dt, t->d.Re); }

// Jacobi sweeps of (1 - dt vis lap / 2) F = R until the maximum residual over the bulk is below
// cnTolerance (ts_jacobi_loop); the reference aborts after (Nx-1)(Ny-1)(Nz-1) sweeps without
// convergence.  F and G swap as the iterates move (G is written everywhere: 0 outside the bulk, as
// the reference's guess field).
static int jacobi(sr_ts *t, const FShape &f, double *&F, double *&G, const double *R, const JacobiCoef &J,
                  const std::vector<GhostWave> &bc, const char *what, int *iters) {
    int per[3] = {t->d.xPer, t->d.yPer, t->d.zPer};
    Bulk B = jacobi_bulk(f, per);
    long maxIterations = (long)(t->N[0] - 1) * (t->N[1] - 1) * (t->N[2] - 1);
    JacobiIO io{t->part, t->dmax, t->hmax, t->jev, t->s};
    int sweeps = 0;
    hipError_t err = hipSuccess;
    int st = ts_jacobi_loop(f, t->M, B, F, G, R, J, bc.data(), (int)bc.size(), io, t->d.cnTolerance, maxIterations,
                            &sweeps, &err);
    HIPCHK(err);
    if (st == 1)
        return set_err(SR_ERR_NOT_CONVERGED, std::string("Jacobi iterations for ") + what +
                                                 " did not converge within the iteration limit");
    if (iters) *iters = sweeps;
    return SR_OK;
}

// hydro_d3::solveVx/Vy/Vz (hydro_d3.cc:370-533), scalar_d3::solveVx/Vy/Vz (scalar_d3.cc:373-531)
static int solve_v(sr_ts *t, int c, int *iters) {
    JacobiCoef J = jacobi_coef(t);
    J.form = t->d.scalar ? 1 : 0;
    J.vis = t->nu;
    for (int k = 0; k < 3; k++) J.ms[k] = t->fv[c].s[k];
    int n = 0;
    RETCHK(jacobi(t, t->fv[c], t->V[c], t->G[c], t->H[c], J, t->vbc[c], c == 0 ? "Vx" : c == 1 ? "Vy" : "Vz", &n));
    t->jac[c] = n;
    if (iters) *iters = n;
    return SR_OK;
}

// scalar_d3::solveT (scalar_d3.cc:533-594): node field; the z metric is the collocated xi_z^2
// array read at node indices, as the reference writes it (scalar_d3.cc:541, 561)
static int solve_t(sr_ts *t, int *iters) {
    JacobiCoef J = jacobi_coef(t);
    J.form = 1;
    J.vis = t->kappa;
    J.ms[0] = 1;
    J.ms[1] = 1;
    J.ms[2] = 0;
    int n = 0;
    RETCHK(jacobi(t, t->fp, t->T, t->Tg, t->Trhs, J, t->tbc, "T", &n));
    t->jac[3] = n;
    if (iters) *iters = n;
    return SR_OK;
}

extern "C" int sr_ts_solve_v(sr_ts *t, int component, int *iters) {
    if (!t || component < 0 || component > 2) return set_err(SR_ERR_ARG, "bad component");
    return solve_v(t, component, iters);
}

extern "C" int sr_ts_divergence(sr_ts *t) {
    if (!t) return set_err(SR_ERR_ARG, "null argument");
    const double *Vc[3] = {t->V[0], t->V[1], t->V[2]};
    ts_divergence(t->fp, t->fv, t->M, Vc, t->divV, t->h, 1.0, 0, t->s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(t->s));
    return SR_OK;
}

extern "C" int sr_ts_energy(sr_ts *t, double *energy) {
    if (!t || !energy) return set_err(SR_ERR_ARG, "null argument");
    const double *Vc[3] = {t->V[0], t->V[1], t->V[2]};
    ts_energy(t->fp, t->fv, t->M, Vc, t->h, t->part, t->dred, t->s);
    HIPCHK(hipMemcpyAsync(t->hred, t->dred, sizeof(double) * 2, hipMemcpyDeviceToHost, t->s));
    HIPCHK(hipStreamSynchronize(t->s));
    HIPCHK(hipGetLastError());
    *energy = t->hred[1] / t->hred[0];
    return SR_OK;
}

extern "C" int sr_ts_max_div(sr_ts *t, double *maxdiv) {
    if (!t || !maxdiv) return set_err(SR_ERR_ARG, "null argument");
    const double *Vc[3] = {t->V[0], t->V[1], t->V[2]};
    ts_divergence(t->fp, t->fv, t->M, Vc, t->divV, t->h, 1.0, 0, t->s);
    ts_max(t->fp, t->divV, t->part, t->dred, t->s);
    HIPCHK(hipMemcpyAsync(t->hred, t->dred, sizeof(double), hipMemcpyDeviceToHost, t->s));
    HIPCHK(hipStreamSynchronize(t->s));
    HIPCHK(hipGetLastError());
    *maxdiv = t->hred[0];
    return SR_OK;
}

// ------------------------------------------------------------------------------------------------
// the step
// ------------------------------------------------------------------------------------------------
extern "C" int sr_ts_time_advance(sr_ts *t, sr_ts_stats *st) {
    if (!t) return set_err(SR_ERR_ARG, "null argument");
    auto t0 = std::chrono::steady_clock::now();
    const double dt = t->d.dt;
    build_rhs(t);
    for (int c = 0; c < 3; c++) RETCHK(solve_v(t, c, nullptr));
    const double *Vc[3] = {t->V[0], t->V[1], t->V[2]};
    // mgRHS = div V / dt (hydro_d3.cc:336-340); pads stay 0
    ts_divergence(t->fp, t->fv, t->M, Vc, t->mgRHS, t->h, 1.0 / dt, 1, t->s);
    HIPCHK(hipGetLastError());
    double guess_ms = 0.0, rel_guess = -1.0, loss = -1.0, rel_linear = -1.0, rel_learned = -1.0, predict_ms = 0.0;
    int choice = -1;
    if (t->guess && t->nsolved >= 1) {
        HIPCHK(hipStreamSynchronize(t->s));
        sr_guess_stats gs;
        RETCHK(sr_guess_predict(t->guess, t->mg, t->Pp1, t->nsolved >= 2 ? t->Pp2 : nullptr, t->mgRHS, t->Pp, &gs));
        choice = gs.choice;
        guess_ms = gs.predict_ms;
        predict_ms = gs.predict_ms;
        rel_linear = gs.rel_linear;
        rel_learned = gs.rel_learned;
        rel_guess = choice == 3 ? gs.rel_linear : choice == 2 ? gs.rel_learned : choice == 1 ? gs.rel_extrap : gs.rel_warm;
    }
    sr_mg_stats ms;
    auto ts0 = std::chrono::steady_clock::now();
    // mgSolve never fails: after vcCount cycles it returns its last iterate and the step goes on
    // (poisson.cc:139-151); a solve that misses its stop rule is reported in the statistics
    int srs = sr_mg_solve(t->mg, t->Pp, t->mgRHS, 1, &ms);
    if (srs != SR_OK && srs != SR_ERR_NOT_CONVERGED) return srs;
    auto ts1 = std::chrono::steady_clock::now();
    t->mg_cycles = ms.cycles;
    // the solver's own evaluation of the guess, where it makes one, is exact (the predictor's fit
    // residual loses accuracy to cancellation below about 1e-7)
    if (choice >= 0 && ms.residual0 >= 0.0) rel_guess = ms.residual0;
    if (t->guess) {
        // the guess's bookkeeping after the solve (training, recording the solved pair) is part of
        // its cost per step; both calls return with their work finished
        auto tg0 = std::chrono::steady_clock::now();
        if (choice >= 0 && t->train) RETCHK(sr_guess_train(t->guess, t->Pp, &loss));
        RETCHK(sr_guess_observe(t->guess, t->mg, t->mgRHS, t->Pp));
        guess_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tg0).count();
    }
    ts_ghost_wave(t->fp, t->Pp, t->sync_p, t->s);
    if (t->guess) {
        // the previous two solutions kept for the guess (its cost too)
        auto tg1 = std::chrono::steady_clock::now();
        std::swap(t->Pp1, t->Pp2);
        HIPCHK(hipMemcpyAsync(t->Pp1, t->Pp, sizeof(double) * fsize(t->fp), hipMemcpyDeviceToDevice, t->s));
        HIPCHK(hipStreamSynchronize(t->s));
        guess_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tg1).count();
    }
    t->nsolved += 1;
    ts_axpy_full(t->P, t->Pp, fsize(t->fp), t->s);
    for (int c = 0; c < 3; c++) ts_project(c, t->fv[c], t->fp, t->M, t->Pp, t->V[c], t->h[c], dt, t->s);
    if (t->d.scalar) {
        // the temperature equation with the projected velocity (scalar_d3.cc:349-363)
        const double *Vp[3] = {t->V[0], t->V[1], t->V[2]};
        ts_build_trhs(t->fp, t->fv, t->M, Vp, t->T, t->Trhs, t->h, t->kappa, dt, t->s);
        ts_ghost_wave(t->fp, t->Trhs, t->sync_p, t->s);
        RETCHK(solve_t(t, nullptr));
    }
    for (int c = 0; c < 3; c++) impose_vbc(t, c);
    if (t->d.scalar) impose_tbc(t);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(t->s));
    auto t1 = std::chrono::steady_clock::now();
    if (st) {
        for (int c = 0; c < 4; c++) st->jacobi_iters[c] = t->jac[c];
        st->solve_ms = std::chrono::duration<double, std::milli>(ts1 - ts0).count();
        st->guess_ms = guess_ms;
        st->sweep0_ms = ms.sweep0_ms;
        st->guess_choice = choice;
        st->rel_guess = rel_guess;
        st->train_loss = loss;
        st->mg_cycles = t->mg_cycles;
        st->mg_converged = ms.converged;
        st->mg_residual = ms.residual;
        st->guess_rel_linear = rel_linear;
        st->guess_rel_learned = rel_learned;
        st->guess_predict_ms = predict_ms;
        st->step_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
    return SR_OK;
}

// ------------------------------------------------------------------------------------------------
// scalar_d3 entry points
// ------------------------------------------------------------------------------------------------
static int need_scalar(const sr_ts *t) {
    if (!t) return set_err(SR_ERR_ARG, "null argument");
    if (!t->T) return set_err(SR_ERR_STATE, "no temperature field (desc.scalar = 0)");
    return SR_OK;
}

extern "C" int sr_ts_coefficients(sr_ts *t, double *nu, double *kappa) {
    if (!t || !nu || !kappa) return set_err(SR_ERR_ARG, "null argument");
    *nu = t->nu;
    *kappa = t->d.scalar ? t->kappa : 0.0;
    return SR_OK;
}

extern "C" int sr_ts_init_rbc(sr_ts *t, unsigned long long seed, double amp) {
    RETCHK(need_scalar(t));
    const long np = fsize(t->fp);
    std::vector<double> T((size_t)np, 0.0);
    uint64_t st = seed;
    for (int i = 0; i < t->fp.n[0]; i++)
        for (int j = 0; j < t->fp.n[1]; j++)
            for (int k = 0; k < t->fp.n[2]; k++) {
                st += 0x9E3779B97F4A7C15ull;
                uint64_t z = st;
                z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
                z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
                z ^= z >> 31;
                double u = ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
                T[(size_t)((((long)(i + 1) * (t->fp.n[1] + 2)) + (j + 1)) * (t->fp.n[2] + 2) + (k + 1))] =
                    amp * (2.0 * u - 1.0);
            }
    HIPCHK(hipMemcpyAsync(t->T, T.data(), sizeof(double) * np, hipMemcpyHostToDevice, t->s));
    ts_fill(t->P, np, 1.0, t->s);
    ts_fill(t->Pp, np, 0.0, t->s);
    for (int c = 0; c < 3; c++) ts_fill(t->V[c], fsize(t->fv[c]), 0.0, t->s);
    for (int c = 0; c < 3; c++) impose_vbc(t, c);
    impose_tbc(t);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(t->s));
    return SR_OK;
}

extern "C" int sr_ts_impose_tbc(sr_ts *t) {
    RETCHK(need_scalar(t));
    impose_tbc(t);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(t->s));
    return SR_OK;
}

extern "C" int sr_ts_build_trhs(sr_ts *t) {
    RETCHK(need_scalar(t));
    const double *Vc[3] = {t->V[0], t->V[1], t->V[2]};
    ts_build_trhs(t->fp, t->fv, t->M, Vc, t->T, t->Trhs, t->h, t->kappa, t->d.dt, t->s);
    ts_ghost_wave(t->fp, t->Trhs, t->sync_p, t->s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(t->s));
    return SR_OK;
}

extern "C" int sr_ts_solve_t(sr_ts *t, int *iters) {
    RETCHK(need_scalar(t));
    return solve_t(t, iters);
}

extern "C" int sr_ts_nusselt(sr_ts *t, double *nu) {
    RETCHK(need_scalar(t));
    if (!nu) return set_err(SR_ERR_ARG, "null argument");
    ts_nusselt(t->fp, t->fv[2], t->M, t->V[2], t->T, t->h, t->part, t->dred, t->s);
    HIPCHK(hipMemcpyAsync(t->hred, t->dred, sizeof(double) * 2, hipMemcpyDeviceToHost, t->s));
    HIPCHK(hipStreamSynchronize(t->s));
    HIPCHK(hipGetLastError());
    *nu = 1.0 + (t->hred[1] / t->hred[0]) / t->kappa;
    return SR_OK;
}

extern "C" int sr_ts_attach_guess(sr_ts *t, sr_guess *g, int train) {
    if (!t) return set_err(SR_ERR_ARG, "null argument");
    t->guess = g;
    t->train = train;
    if (g && !t->Pp1) {
        const long np = fsize(t->fp);
        HIPCHK(hipMalloc((void **)&t->Pp1, sizeof(double) * np));
        HIPCHK(hipMalloc((void **)&t->Pp2, sizeof(double) * np));
        ts_fill(t->Pp1, np, 0.0, t->s);
        ts_fill(t->Pp2, np, 0.0, t->s);
        t->nsolved = 0;
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(t->s));
    }
    return SR_OK;
}

// ------------------------------------------------------------------------------------------------
// measurement
// ------------------------------------------------------------------------------------------------
static Bulk bulk_of(const sr_ts *t, const FShape &f) {
    int per[3] = {t->d.xPer, t->d.yPer, t->d.zPer};
    Bulk B{};
    for (int k = 0; k < 3; k++) {
        B.lo[k] = 0;
        B.hi[k] = f.n[k] - 1;
        if (f.s[k] && !per[k]) {
            B.lo[k] = 1;
            B.hi[k] = f.n[k] - 2;
        }
    }
    return B;
}

static long core_n(const FShape &f) { return (long)f.n[0] * f.n[1] * f.n[2]; }

extern "C" int sr_ts_kernel_bytes(const sr_ts *t, int kernel, double *bytes) {
    if (!t || !bytes) return set_err(SR_ERR_ARG, "null argument");
    const double w = 8.0;
    double b = 0.0;
    switch (kernel) {
        case 0:   // nseRHS of the three components: V_c, the two other components' slices, P read, H_c written
            for (int c = 0; c < 3; c++) b += 5.0 * w * (double)core_n(t->fv[c]);
            break;
        case 1: b = 3.0 * w * (double)fsize(t->fv[0]); break;           // Jacobi sweep of Vx: V, R read, G written
        case 2: {                                                       // its residual over the bulk: V, R read
            Bulk B = bulk_of(t, t->fv[0]);
            b = 2.0 * w * (double)(B.hi[0] - B.lo[0] + 1) * (B.hi[1] - B.lo[1] + 1) * (B.hi[2] - B.lo[2] + 1);
            break;
        }
        case 3: b = 4.0 * w * (double)core_n(t->fp); break;            // divergence: three faces read, div written
        case 4:                                                         // projection: Pp read, V_c read and written
            for (int c = 0; c < 3; c++) b += 3.0 * w * (double)core_n(t->fv[c]);
            break;
        case 5: b = 3.0 * w * (double)core_n(t->fp); break;            // energy: three faces read
        case 6: b = 3.0 * w * (double)fsize(t->fv[0]); break;           // Jacobi sweep of Vx with the residual of
                                                                        // its input: V, R read, G written
        default: return set_err(SR_ERR_ARG, "unknown kernel id");
    }
    *bytes = b;
    return SR_OK;
}

extern "C" int sr_ts_time_kernel(sr_ts *t, int kernel, int reps, double *avg_ms) {
    if (!t || !avg_ms || reps < 1 || kernel < 0 || kernel > 6) return set_err(SR_ERR_ARG, "bad arguments");
    const double *Vc[3] = {t->V[0], t->V[1], t->V[2]};
    const double visc = t->d.scalar ? t->nu : 1.0 / t->d.Re;
    JacobiCoef J = jacobi_coef(t);
    J.form = t->d.scalar ? 1 : 0;
    J.vis = t->nu;
    for (int k = 0; k < 3; k++) J.ms[k] = t->fv[0].s[k];
    Bulk B = bulk_of(t, t->fv[0]);
    auto launch = [&]() {
        switch (kernel) {
            case 0:
                for (int c = 0; c < 3; c++)
                    ts_build_rhs(c, t->fv, t->fp, t->M, Vc, t->P, t->H[c], t->h, visc, t->d.dt,
                                 t->d.scalar ? t->T : nullptr, t->Fb, t->s);
                break;
            case 1: ts_jacobi(t->fv[0], t->M, B, t->V[0], t->H[0], t->G[0], J, t->s); break;
            case 2: ts_jacobi_res(t->fv[0], t->M, B, t->V[0], t->H[0], J, t->part, t->dmax, t->s); break;
            case 3: ts_divergence(t->fp, t->fv, t->M, Vc, t->divV, t->h, 1.0, 0, t->s); break;
            case 4:
                for (int c = 0; c < 3; c++) ts_project(c, t->fv[c], t->fp, t->M, t->Pp, t->V[c], t->h[c], t->d.dt, t->s);
                break;
            case 5: ts_energy(t->fp, t->fv, t->M, Vc, t->h, t->part, t->dred, t->s); break;
            case 6: ts_jacobi_vres(t->fv[0], t->M, B, t->V[0], t->H[0], t->G[0], J, t->part, t->dmax, t->s); break;
        }
    };
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    launch();
    HIPCHK(hipEventRecord(e0, t->s));
    for (int r = 0; r < reps; r++) launch();
    HIPCHK(hipEventRecord(e1, t->s));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    hipError_t e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(e);
    HIPCHK(hipGetLastError());
    *avg_ms = (double)ms / reps;
    return SR_OK;
}

extern "C" int sr_ts_pressure_solver(sr_ts *t, sr_mg **mg) {
    if (!t || !mg) return set_err(SR_ERR_ARG, "null argument");
    *mg = t->mg;
    return SR_OK;
}
