// Solve drivers of the C ABI: poisson::mgSolve (lib/poisson/poisson.cc:100-151) on the device,
// in FP64 and in mixed precision, plus the kernel-timing entry points used for roofline figures.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/saras_mg.h"
#include "mg_kernels.h"
#include "mg_state.h"

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

// the 32-byte read-back of the cycle's reduced quantities into the pinned host mirror; a failure
// is kept in m->body_err (the call may be recorded into a graph, where it cannot return one)
static void read_back(sr_mg *m) {
    hipError_t e = hipMemcpyAsync(m->hres, m->dres, sizeof(double) * 4, hipMemcpyDeviceToHost, m->stream);
    if (e != hipSuccess && m->body_err == hipSuccess) m->body_err = e;
}

// before every cycle: the host mirror holds NaN until the cycle's read-back lands, so a read-back
// that did not happen shows as a solve that never meets its stop rule, not as a stale norm
static void poison_hres(sr_mg *m) {
    for (int i = 0; i < 4; i++) m->hres[i] = std::nan("");
}

static bool is_singular(const sr_mg *m) {
    return !(m->d.testPoisson && !(m->d.xPer && m->d.yPer && m->d.zPer));
}

static Bc pads_rules(const sr_mg *m) {
    Bc b{};
    b.rule[0] = m->d.xPer ? G_PERIODIC : G_KEEP;
    b.rule[1] = m->d.yPer ? G_PERIODIC : G_KEEP;
    b.rule[2] = G_KEEP;
    return b;
}

static int stage_alloc(sr_mg *m) {
    if (m->stage) return SR_OK;
    const Lay &L = m->lv[0].L;
    size_t n = (size_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    HIPCHK(hipMalloc((void **)&m->stage, sizeof(double) * n));
    HIPCHK(hipMalloc((void **)&m->stage2, sizeof(double) * n));
    return SR_OK;
}

// Loads the right-hand side and the initial guess into level 0 and applies the solvability step:
// reference semantics subtract the mean of the padded RHS (poisson.cc:120-129); projected
// semantics remove the component along the left null vector of L.
static void load_level0(sr_mg *m, const double *drhs, const double *dp) {
    Level &V = m->lv[0];
    hipStream_t s = m->stream;
    bool proj = m->d.semantics == SR_SEM_PROJECTED && is_singular(m);
    if (proj) {
        // the weighted sum read from the caller's array, then one pass that packs it with the shift
        // and reduces the computeError terms of the result (pads_rules: x/y periodic images and z
        // ghosts kept at 0, so the padded maximum is the core one)
        launch_wsum_ref(V.L, drhs, m->wvec[0], m->wvec[1], m->wvec[2], m->part, m->dres, s);
        launch_pack_shift(V.L, drhs, V.f, m->dres, -1.0 / m->wsum, m->part, m->fterms, s);
    } else {
        launch_pack_full<double, double>(V.L, drhs, V.f, s);
    }
    launch_pack_full<double, double>(V.L, dp, V.u, s);
    launch_fill_ghosts<double>(V.L, V.f, pads_rules(m), s);
    launch_fill_ghosts<double>(V.L, V.u, pads_rules(m), s);
    if (m->d.semantics == SR_SEM_REFERENCE && !m->d.testPoisson) {
        double npad = (double)(V.L.nx + 2) * (V.L.ny + 2) * (V.L.nz + 2);
        launch_rhs_sums<double>(V.L, V.f, nullptr, nullptr, nullptr, m->part, m->dres, s);
        launch_add_scalar<double>(V.L, V.f, m->dres + 1, -1.0 / npad, 0, s);
    }
    if (!proj) f_terms_level0(m);
}

static int norm_order(const sr_mg *m) {
    return m->d.semantics == SR_SEM_PROJECTED ? 2 : m->d.resType;
}

// one FP64 cycle plus its norm and the read-back of the reduced quantities
static void cycle_fp64(sr_mg *m) {
    vcycle_fp64(m);
    // a Gauss-Seidel cycle ends on a colour-1 half sweep of level 0, which leaves the colour-1
    // residual zero (up to the rounding of f - Lu): the norm is taken over colour 0
    norm_level0(m, m->d.gsSmooth && m->d.postSmooth > 0 ? 0 : -1);
    read_back(m);
}

static bool graph_ok(const sr_mg *m) {
    // Jacobi swaps the level handles; the graph replays fixed pointers, so odd counts run eagerly
    return m->d.gsSmooth || ((m->d.preSmooth % 2 == 0) && (m->d.postSmooth % 2 == 0));
}

// Capture `body` into an executable graph.  A capture the runtime rejects is retried once without
// the timing marks; if that fails too the cycle runs eagerly from then on.
static bool capture(sr_mg *m, void (*body)(sr_mg *), hipGraph_t &g, hipGraphExec_t &ex) {
    for (int attempt = 0; attempt < 2; attempt++) {
        g = nullptr;
        ex = nullptr;
        if (hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) break;
        m->body_err = hipSuccess;
        body(m);
        hipError_t e1 = hipStreamEndCapture(m->stream, &g);
        bool recorded = m->body_err == hipSuccess;
        m->body_err = hipSuccess;
        if (e1 == hipSuccess && recorded && g && hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) == hipSuccess)
            return true;
        (void)hipGetLastError();
        if (g) (void)hipGraphDestroy(g);
        g = nullptr;
        ex = nullptr;
        m->marks_off = true;
    }
    (void)hipGetLastError();
    m->graphs_off = true;
    return false;
}

// a body run eagerly: its launches' and its read-back's status
static int run_eager(sr_mg *m, void (*body)(sr_mg *)) {
    m->body_err = hipSuccess;
    body(m);
    hipError_t e = m->body_err;
    m->body_err = hipSuccess;
    HIPCHK(e);
    HIPCHK(hipGetLastError());
    return SR_OK;
}

static int run_cycle(sr_mg *m, int kind, void (*body)(sr_mg *)) {
    poison_hres(m);
    if (!graph_ok(m) || m->graphs_off) return run_eager(m, body);
    if (m->graph_kind != kind) {
        if (m->graph_exec) { (void)hipGraphExecDestroy(m->graph_exec); m->graph_exec = nullptr; }
        if (m->graph) { (void)hipGraphDestroy(m->graph); m->graph = nullptr; }
        if (!capture(m, body, m->graph, m->graph_exec)) return run_eager(m, body);
        m->graph_kind = kind;
    }
    HIPCHK(hipGraphLaunch(m->graph_exec, m->stream));
    return SR_OK;
}

static int solve_fp64(sr_mg *m, int &cycles, double &res, bool &conv, double &sweep_ms) {
    int nt = norm_order(m);
    double prev = 0.0;
    conv = false;
    double tsum = 0.0;
    int tcount = 0;
    for (int it = 0; it < m->d.vcCount; it++) {
        RETCHK(run_cycle(m, 0, cycle_fp64));
        HIPCHK(hipStreamSynchronize(m->stream));
        if (m->d.gsSmooth && m->d.preSmooth > 0) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, m->sw0, m->sw1) == hipSuccess) {
                tsum += ms;
                tcount += 2 * m->d.preSmooth;
            }
        }
        double r = error_from(m, m->hres, nt);
        m->history.push_back(r);
        cycles = it + 1;
        res = r;
        if (m->d.semantics == SR_SEM_PROJECTED) {
            if (r < m->d.relTolerance) { conv = true; break; }
        } else if (!m->d.testPoisson) {
            if (std::fabs(prev - r) < m->d.mgTolerance) { conv = true; break; }
            prev = r;
        }
    }
    if (m->d.testPoisson && m->d.semantics != SR_SEM_PROJECTED) conv = true;   // fixed cycle count
    sweep_ms = tcount ? tsum / tcount : 0.0;
    return SR_OK;
}

// ------------------------------------------------------------------------------------------------
// mixed precision: u (FP64, two ping-pong buffers: level-0 u and the unused level-0 tmp) is
// corrected by FP32 V-cycles on the error equation; residual, correction and norm stay FP64
// ------------------------------------------------------------------------------------------------

static void mixed_body_ab(sr_mg *m) {
    Level &V = m->lv[0];
    vcycle_fp32(m);
    launch_outer_residual(V.L, V.u, V.u32, V.f, V.t, V.f32, V.c64[0], V.c64[1], V.c64[2], pressure_bc(m, 0, 0),
                          m->part2, m->fterms, m->dres, m->stream);
    read_back(m);
}

static void mixed_body_ba(sr_mg *m) {
    Level &V = m->lv[0];
    vcycle_fp32(m);
    launch_outer_residual(V.L, V.t, V.u32, V.f, V.u, V.f32, V.c64[0], V.c64[1], V.c64[2], pressure_bc(m, 0, 0),
                          m->part2, m->fterms, m->dres, m->stream);
    read_back(m);
}

static int run_graph2(sr_mg *m, int which) {
    void (*body)(sr_mg *) = which == 0 ? mixed_body_ab : mixed_body_ba;
    hipGraphExec_t &ex = m->graph_mixed[which];
    poison_hres(m);
    if (!ex && !m->graphs_off) capture(m, body, m->graph_mixed_g[which], ex);
    if (!ex) return run_eager(m, body);
    HIPCHK(hipGraphLaunch(ex, m->stream));
    return SR_OK;
}

static int solve_mixed(sr_mg *m, int &cycles, double &res, bool &conv, double &sweep_ms, double *&dsol) {
    if (!m->d.gsSmooth) return set_err(SR_ERR_ARG, "mixed precision needs gsSmooth = 1");
    RETCHK(ensure_fp32(m));
    Level &V = m->lv[0];
    long sz = lay_size(V.L);
    // ghosts of the initial guess under its own BCs (Dirichlet walls included), then the second
    // buffer of the pair inherits them; wall ghosts are never rewritten afterwards
    launch_fill_ghosts<double>(V.L, V.u, pressure_bc(m, 0, 0), m->stream);
    HIPCHK(hipMemcpyAsync(V.t, V.u, sizeof(double) * sz, hipMemcpyDeviceToDevice, m->stream));
    launch_outer_residual(V.L, V.u, nullptr, V.f, nullptr, V.f32, V.c64[0], V.c64[1], V.c64[2],
                          pressure_bc(m, 0, 0), m->part2, m->fterms, m->dres, m->stream);
    HIPCHK(hipMemcpyAsync(m->hres, m->dres, sizeof(double) * 4, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    int nt = norm_order(m);
    double r = error_from(m, m->hres, nt);
    dsol = V.u;
    conv = false;
    cycles = 0;
    res = r;
    m->res0 = r;
    if (m->d.semantics == SR_SEM_PROJECTED && r < m->d.relTolerance) { conv = true; return SR_OK; }
    double prev = 0.0, tsum = 0.0;
    int tcount = 0;
    int cur = 0;
    for (int it = 0; it < m->d.vcCount; it++) {
        RETCHK(run_graph2(m, cur));
        HIPCHK(hipStreamSynchronize(m->stream));
        float ms = 0.f;
        if (m->pre32 > 0 && hipEventElapsedTime(&ms, m->sw0, m->sw1) == hipSuccess) {
            tsum += ms;
            tcount += 2 * m->pre32;
        }
        cur ^= 1;
        dsol = cur ? V.t : V.u;
        r = error_from(m, m->hres, nt);
        m->history.push_back(r);
        cycles = it + 1;
        res = r;
        if (m->d.semantics == SR_SEM_PROJECTED) {
            if (r < m->d.relTolerance) { conv = true; break; }
        } else if (!m->d.testPoisson) {
            if (std::fabs(prev - r) < m->d.mgTolerance) { conv = true; break; }
            prev = r;
        }
    }
    if (m->d.testPoisson && m->d.semantics != SR_SEM_PROJECTED) conv = true;
    sweep_ms = tcount ? tsum / tcount : 0.0;
    return SR_OK;
}

// ------------------------------------------------------------------------------------------------
// Krylov-accelerated mixed precision: flexible GCR(1) with one FP32 V-cycle as the preconditioner.
// The iterate ping-pongs between the level-0 u and tmp buffers like the refinement loop, the FP32
// search direction between zdir[0] and zdir[1]; the step coefficients stay on the device, so one
// iteration is one captured graph with a single read-back of the residual norms.
// ------------------------------------------------------------------------------------------------
static void gcr_body(sr_mg *m, int cur) {
    Level &V = m->lv[0];
    hipStream_t s = m->stream;
    double *x_in = cur ? V.t : V.u, *x_out = cur ? V.u : V.t;
    const float *z = m->zdir[cur];
    float *z_next = m->zdir[cur ^ 1];
    vcycle_fp32(m);
    // the FP32 cycle ends on level 0's post-smoothing (a colour-1 half sweep last) when level 0 is
    // above the bottom of the cycle and has post-smoothing sweeps
    const int Lc = m->coarse_enable ? m->coarse_first32 : (1 << 20);
    const bool q1r = std::min(Lc, m->depth) >= 1 && m->post32 > 0;
    launch_gcr_dots(V.L, V.u32, z, V.f32, V.c32[0], V.c32[1], V.c32[2], m->gpart, m->gcoef, q1r, s,
                    &m->fb_coef);
    launch_gcr_step(V.L, x_in, V.u32, z, m->gcoef, x_out, z_next, pressure_bc(m, 0, 0), error_bc(m, 0, 0), s);
    launch_outer_residual(V.L, x_out, nullptr, V.f, nullptr, V.f32, V.c64[0], V.c64[1], V.c64[2],
                          pressure_bc(m, 0, 0), m->part2, m->fterms, m->dres, s);
    read_back(m);
}
static void gcr_body_0(sr_mg *m) { gcr_body(m, 0); }
static void gcr_body_1(sr_mg *m) { gcr_body(m, 1); }

static int run_gcr_graph(sr_mg *m, int which) {
    void (*body)(sr_mg *) = which == 0 ? gcr_body_0 : gcr_body_1;
    hipGraphExec_t &ex = m->graph_gcr[which];
    poison_hres(m);
    if (!ex && !m->graphs_off) capture(m, body, m->graph_gcr_g[which], ex);
    if (!ex) return run_eager(m, body);
    HIPCHK(hipGraphLaunch(ex, m->stream));
    return SR_OK;
}

static int solve_gcr(sr_mg *m, int &cycles, double &res, bool &conv, double &sweep_ms, double *&dsol) {
    if (!m->d.gsSmooth) return set_err(SR_ERR_ARG, "mixed precision needs gsSmooth = 1");
    Level &V = m->lv[0];
    long sz = lay_size(V.L);
    hipStream_t s = m->stream;
    Bc bc0 = pressure_bc(m, 0, 0);
    launch_fill_ghosts<double>(V.L, V.u, bc0, s);
    // the second iterate buffer needs the fixed wall ghosts only; every other face ghost is
    // rewritten as an image by each update
    if (bc0.rule[0] == G_WALL || bc0.rule[1] == G_WALL || bc0.rule[2] == G_WALL)
        HIPCHK(hipMemcpyAsync(V.t, V.u, sizeof(double) * sz, hipMemcpyDeviceToDevice, s));
    // no previous direction: the first step is the minimal-residual step along the cycle's result
    launch_fill<float>(m->zdir[0], sz, 0.0f, s);
    launch_outer_residual(V.L, V.u, nullptr, V.f, nullptr, V.f32, V.c64[0], V.c64[1], V.c64[2],
                          pressure_bc(m, 0, 0), m->part2, m->fterms, m->dres, s);
    HIPCHK(hipMemcpyAsync(m->hres, m->dres, sizeof(double) * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int nt = norm_order(m);
    double r = error_from(m, m->hres, nt);
    dsol = V.u;
    conv = false;
    cycles = 0;
    res = r;
    m->res0 = r;
    if (m->d.semantics == SR_SEM_PROJECTED && r < m->d.relTolerance) { conv = true; return SR_OK; }
    double prev = 0.0, tsum = 0.0;
    int tcount = 0;
    int cur = 0;
    for (int it = 0; it < m->d.vcCount; it++) {
        RETCHK(run_gcr_graph(m, cur));
        HIPCHK(hipStreamSynchronize(s));
        float ms = 0.f;
        if (m->pre32 > 0 && hipEventElapsedTime(&ms, m->sw0, m->sw1) == hipSuccess) {
            tsum += ms;
            tcount += 2 * m->pre32;
        }
        cur ^= 1;
        dsol = cur ? V.t : V.u;
        r = error_from(m, m->hres, nt);
        m->history.push_back(r);
        cycles = it + 1;
        res = r;
        if (m->d.semantics == SR_SEM_PROJECTED) {
            if (r < m->d.relTolerance) { conv = true; break; }
        } else if (!m->d.testPoisson) {
            if (std::fabs(prev - r) < m->d.mgTolerance) { conv = true; break; }
            prev = r;
        }
    }
    if (m->d.testPoisson && m->d.semantics != SR_SEM_PROJECTED) conv = true;
    sweep_ms = tcount ? tsum / tcount : 0.0;
    return SR_OK;
}

// computeError (poisson_d3.cc:391-459) of an arbitrary guess: the right-hand side and the guess
// are loaded as sr_mg_solve loads them (projection or mean removal included), the guess's ghosts
// filled, and the stop rule's relative residual returned; optionally f - L p in padded layout.
extern "C" int sr_mg_residual(sr_mg *m, const double *p, const double *rhs, double *r, double *rel) {
    if (!m || !p || !rhs) return set_err(SR_ERR_ARG, "null argument");
    Level &V = m->lv[0];
    load_level0(m, rhs, p);
    launch_fill_ghosts<double>(V.L, V.u, pressure_bc(m, 0, 0), m->stream);
    norm_level0(m);
    HIPCHK(hipMemcpyAsync(m->hres, m->dres, sizeof(double) * 4, hipMemcpyDeviceToHost, m->stream));
    if (r) {
        residual_level(m, 0);
        launch_unpack<double, double>(V.L, V.t, r, m->stream);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(m->stream));
    if (rel) *rel = error_from(m, m->hres, norm_order(m));
    return SR_OK;
}

extern "C" int sr_mg_solve(sr_mg *m, double *p, const double *rhs, int on_device, sr_mg_stats *st) {
    if (!m || !p || !rhs) return set_err(SR_ERR_ARG, "null argument");
    auto t0 = std::chrono::steady_clock::now();
    const Lay &L = m->lv[0].L;
    size_t n = (size_t)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    const double *drhs = rhs;
    double *dp = p;
    if (!on_device) {
        RETCHK(stage_alloc(m));
        HIPCHK(hipMemcpyAsync(m->stage, rhs, sizeof(double) * n, hipMemcpyHostToDevice, m->stream));
        HIPCHK(hipMemcpyAsync(m->stage2, p, sizeof(double) * n, hipMemcpyHostToDevice, m->stream));
        drhs = m->stage;
        dp = m->stage2;
    }
    m->history.clear();
    m->vLevel = 0;
    load_level0(m, drhs, dp);
    HIPCHK(hipGetLastError());
    int cycles = 0;
    double res = 0.0;
    m->res0 = -1.0;
    bool conv = false;
    double sweep_ms = 0.0;
    Level &V = m->lv[0];
    double *dsol = V.u;
    if (m->d.precision == SR_PREC_FP64) {
        RETCHK(solve_fp64(m, cycles, res, conv, sweep_ms));
    } else if (m->d.precision == SR_PREC_MIXED) {
        RETCHK(solve_mixed(m, cycles, res, conv, sweep_ms, dsol));
    } else if (m->d.precision == SR_PREC_MIXED_GCR) {
        RETCHK(solve_gcr(m, cycles, res, conv, sweep_ms, dsol));
    } else {
        return set_err(SR_ERR_ARG, "unknown precision mode");
    }
    // closing ghosts of the last post-smooth (poisson_d3.cc:181) and the copy over the padded inFn
    // (poisson.cc:151)
    launch_fill_ghosts<double>(V.L, dsol, pressure_bc(m, 0, 0), m->stream);
    launch_unpack<double, double>(V.L, dsol, dp, m->stream);
    if (!on_device) HIPCHK(hipMemcpyAsync(p, dp, sizeof(double) * n, hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipStreamSynchronize(m->stream));
    HIPCHK(hipGetLastError());
    m->vLevel = 0;
    m->zeroBC = 0;
    auto t1 = std::chrono::steady_clock::now();
    if (st) {
        st->cycles = cycles;
        st->converged = conv ? 1 : 0;
        st->residual = res;
        st->residual0 = m->res0;
        st->solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st->sweep0_ms = sweep_ms;
    }
    if (!conv && m->d.semantics == SR_SEM_PROJECTED)
        return set_err(SR_ERR_NOT_CONVERGED, "relative residual above relTolerance after vcCount cycles");
    return SR_OK;
}

// ------------------------------------------------------------------------------------------------
// measurement
// ------------------------------------------------------------------------------------------------
extern "C" int sr_mg_kernel_bytes(const sr_mg *m, int kernel, int level, double *bytes) {
    if (!m || !bytes || level < 0 || level > m->depth) return set_err(SR_ERR_ARG, "bad arguments");
    const Lay &L = m->lv[level].L;
    double nn = (double)L.nx * L.ny * L.nz;
    double w = 8.0;
    switch (kernel) {
        case 0: *bytes = 12.0 * nn / 8.0 * w; break;              // half sweep: u(other), f, u(this) halves
        case 1: *bytes = 3.0 * nn * w; break;                     // residual: read u, f; write t
        case 2: {                                                  // restriction: read fine t, write coarse f
            if (level >= m->depth) return set_err(SR_ERR_ARG, "no coarser level");
            const Lay &Cl = m->lv[level + 1].L;
            *bytes = (nn + (double)Cl.nx * Cl.ny * Cl.nz) * w;
            break;
        }
        case 3: {                                                  // prolongation-correction into `level`
            if (level >= m->depth) return set_err(SR_ERR_ARG, "no coarser level");
            const Lay &Cl = m->lv[level + 1].L;
            *bytes = (2.0 * nn + (double)Cl.nx * Cl.ny * Cl.nz) * w;
            break;
        }
        case 4: {
            double g = (double)(L.nx + 2) * (L.ny + 2) * (L.nz + 2) - nn;
            *bytes = 2.0 * g * w;
            break;
        }
        case 5: *bytes = 2.0 * nn * w; break;                     // norm: read u, f
        case 6: *bytes = 12.0 * nn / 8.0 * 4.0; break;            // FP32 half sweep
        case 7:
        case 8: *bytes = 0.0; break;                              // coarse cycle: latency-bound, no HBM figure
        default: return set_err(SR_ERR_ARG, "unknown kernel id");
    }
    return SR_OK;
}

extern "C" int sr_mg_time_kernel(sr_mg *m, int kernel, int level, int reps, double *avg_ms) {
    if (!m || !avg_ms || reps < 1 || level < 0 || level > m->depth) return set_err(SR_ERR_ARG, "bad arguments");
    if ((kernel == 2 || kernel == 3) && level >= m->depth) return set_err(SR_ERR_ARG, "no coarser level");
    Level &V = m->lv[level];
    hipStream_t s = m->stream;
    Bc bc = pressure_bc(m, level, 1);
    auto launch = [&](int r) {
        switch (kernel) {
            case 0: launch_rb_sweep<double>(V.L, V.u, V.f, V.c64[0], V.c64[1], V.c64[2], r & 1, bc, false, s); break;
            case 1: launch_residual<double, double>(V.L, V.u, V.f, V.t, V.c64[0], V.c64[1], V.c64[2], s); break;
            case 2: launch_restrict<double>(V.L, m->lv[level + 1].L, V.t, m->lv[level + 1].f, s); break;
            case 3: launch_prolong<double, double>(V.L, m->lv[level + 1].L, m->lv[level + 1].u, V.u, 1, s); break;
            case 4: launch_fill_ghosts<double>(V.L, V.u, bc, s); break;
            case 5:
                launch_norm_res<double>(V.L, V.u, V.f, V.c64[0], V.c64[1], V.c64[2], m->part2, m->fterms, m->dres, s);
                break;
            case 6:
                launch_rb_sweep<float>(V.L, V.u32, V.f32, V.c32[0], V.c32[1], V.c32[2], r & 1, bc, false, s);
                break;
            case 7: {
                Level &C = m->lv[m->coarse_first64];
                launch_coarse_cycle<double>(m->cdesc64, m->clds64, C.u, C.f, m->diters, s);
                break;
            }
            case 8: {
                Level &C = m->lv[m->coarse_first32];
                launch_coarse_cycle<float>(m->cdesc32, m->clds32, C.u32, C.f32, m->diters, s);
                break;
            }
        }
    };
    if (kernel < 0 || kernel > 8) return set_err(SR_ERR_ARG, "unknown kernel id");
    if ((kernel == 6 || kernel == 8) && !V.u32)
        return set_err(SR_ERR_STATE, "no FP32 hierarchy (mixed precision not enabled)");
    if (kernel == 7 && m->coarse_first64 > m->depth) return set_err(SR_ERR_STATE, "no single-workgroup coarse cycle");
    if (kernel == 8 && m->coarse_first32 > m->depth) return set_err(SR_ERR_STATE, "no single-workgroup coarse cycle");
    launch(0);
    launch(1);
    HIPCHK(hipEventRecord(m->ev0, s));
    for (int r = 0; r < reps; r++) launch(r);
    HIPCHK(hipEventRecord(m->ev1, s));
    HIPCHK(hipEventSynchronize(m->ev1));
    HIPCHK(hipGetLastError());
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, m->ev0, m->ev1));
    *avg_ms = (double)ms / reps;
    return SR_OK;
}
