// Internal state of a solver handle (include/saras_mg.h) and the operators shared between the
// per-operator entry points (mg_api.hip) and the solve drivers (mg_solve.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/saras_mg.h"
#include "mg_layout.h"

namespace sr {

struct Level {
    Lay L{};
    double h[3] = {0, 0, 0};              // hx(l), hy(l), hz(l)
    std::vector<double> coef_h[3];        // host copy: per padded index {lp, lm, lc, 0}
    double *c64[3] = {nullptr, nullptr, nullptr};
    float *c32[3] = {nullptr, nullptr, nullptr};
    double *u = nullptr, *f = nullptr, *t = nullptr;        // FP64 pressureData / residualData / tmp
    float *u32 = nullptr, *f32 = nullptr, *t32 = nullptr;   // FP32 hierarchy (mixed precision)
    double *wc[3] = {nullptr, nullptr, nullptr};            // null vectors of this level (coarsest)
    double wcsum = 0;
};

}  // namespace sr

struct sr_mg {
    sr_mg_desc d{};
    int depth = 0;
    int vLevel = 0;
    int zeroBC = 0;
    int maxCount = 0;
    std::vector<sr::Level> lv;
    std::vector<double> met0[6];
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t sw0 = nullptr, sw1 = nullptr;   // around the finest pre-smoothing sweeps of a cycle
    bool marks_off = false;
    hipError_t body_err = hipSuccess;   // first failed call inside a cycle body (captured or eager)
    bool graphs_off = false;
    double *wall[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    double *wvec[3] = {nullptr, nullptr, nullptr};
    double wsum = 0;
    double *part = nullptr;      // reduction partials
    double *part2 = nullptr;     // per-block (sum, max) partials of the level-0 residual norm
    double *fterms = nullptr;    // right-hand side terms of computeError (device, 4 doubles)
    double *fstage = nullptr;    // stage of the GCR coefficients' two-stage final (FINAL_STAGE doubles)
    unsigned *fcount = nullptr;  // its counter (zero between launches)
    sr::FinalBufs fb_coef{};
    double *dres = nullptr;      // reduced quantities (device)
    double *hres = nullptr;      // pinned host mirror
    int *diters = nullptr;
    double *stage = nullptr, *stage2 = nullptr;   // reference-layout staging (finest level)
    float *r32 = nullptr;        // FP32 finest residual (mixed precision)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_kind = -1;
    // one-workgroup coarse end of the cycle (levels coarse_first..depth in LDS); first > depth = off
    int coarse_first64 = 1 << 20, coarse_first32 = 1 << 20;
    sr::CoarseDesc cdesc64{}, cdesc32{};
    size_t clds64 = 0, clds32 = 0;
    int coarse_enable = 1;
    long lds_limit = 0;
    hipGraph_t graph_mixed_g[2] = {nullptr, nullptr};
    hipGraphExec_t graph_mixed[2] = {nullptr, nullptr};
    // Krylov-accelerated mixed precision: FP32 search directions (ping-pong), block sums, step
    // coefficients and the two captured iteration bodies
    // exact coarsest solve (coarseDirect): dense inverse, transposed, FP64 and FP32 copies
    double *pinv64 = nullptr;
    float *pinv32 = nullptr;
    int ncd = 0;
    float *zdir[2] = {nullptr, nullptr};
    int pre32 = 0, post32 = 0;  // smoothing counts of the FP32 cycle (innerPre/PostSmooth or the outer ones)
    double res0 = -1.0;        // stop-rule residual of the last solve's initial guess (-1: not evaluated)
    double *gpart = nullptr;
    double *gcoef = nullptr;
    hipGraph_t graph_gcr_g[2] = {nullptr, nullptr};
    hipGraphExec_t graph_gcr[2] = {nullptr, nullptr};
    std::vector<double> history;
};

namespace sr {

int set_err(int code, const std::string &msg);
void mark(sr_mg *m, hipEvent_t e);
Bc pressure_bc(const sr_mg *m, int level, int zeroBC);
Bc tmp_bc(const sr_mg *m);
// ghost rules of a correction (pressure rules with homogeneous walls)
Bc error_bc(const sr_mg *m, int l, int zeroBC);
// faces_ready: the level's face-ghost images are current (written by the prolongation), so the
// red-black sweeps start without a ghost fill
// first_done: the first (zero-start) colour-0 half sweep was done by the level's restriction
void smooth_level(sr_mg *m, int l, int count, int zeroBC, bool zero_start, bool exact,
                  hipEvent_t ea = nullptr, hipEvent_t eb = nullptr, bool faces_ready = false,
                  bool first_done = false);
// images: with color 0 and image residual-pad rules the kernel writes the face and edge ghosts the
// red-black restriction reads instead of a separate fill
void residual_level(sr_mg *m, int l, int color = -1, bool images = false);
void coarsest_level(sr_mg *m, int l, int zeroBC);
void vcycle_fp64(sr_mg *m);
void vcycle_fp32(sr_mg *m);
int ensure_fp32(sr_mg *m);
void norm_level0(sr_mg *m, int color = -1);   // needs the right-hand side terms of f_terms_level0
void f_terms_level0(sr_mg *m);
double error_from(const sr_mg *m, const double *q, int norm);

template <typename T>
void launch_jacobi(const Lay &L, const T *u, const T *f, T *t, const T *cx, const T *cy, const T *cz,
                   hipStream_t s);
template <typename T>
void launch_project_small(const Lay &L, T *f, const double *wx, const double *wy, const double *wz, double wsum,
                          hipStream_t s);

}  // namespace sr
