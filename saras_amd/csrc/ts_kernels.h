// Kernel-side types and launchers of the hydro time step (ts_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace sr {

// core sizes of a padded field and its staggering per direction (1 node, 0 face)
struct FShape {
    int n[3];
    int s[3];
};

// device metric arrays: m[dim][0 node / 1 face][quantity][index + 1];
// quantity 0 coordinate, 1 xi_x, 2 xi_xx, 3 xi_x^2
struct Metrics {
    const double *m[3][2][4];
};

// inclusive index box (bulk of a field, field.cc:206-219)
struct Bulk {
    int lo[3], hi[3];
};

// scalars of the Jacobi viscous / diffusive solves (hydro_d3.cc:375-410, scalar_d3.cc:378-410,
// 538-570).  form 0: hydro_d3's 1/Re expressions; form 1: scalar_d3's nu (or kappa) expressions.
// ms[d]: metric family of direction d (1 node, 0 face), read at the field's own indices.
struct JacobiCoef {
    double hx, hy, hz;
    double hx2hy2, hy2hz2, hz2hx2, hx2hy2hz2;
    double dt, Re, inverseRe;
    double vis;
    int form;
    int ms[3];
    // set by the launchers from the fields above: the grid-uniform quotients of the sweep and of the
    // residual, so that a node divides once (the sweep's per-node denominator) instead of six times
    double rA, rD;            // sweep: G = (a rA + R) / (1 + d rD)
    double ih2[3], lapS;      // residual: V - lapS sum_d m_d (V(+1) - 2 V + V(-1)) ih2_d
};

// one ghost / wall assignment F(dst) = a F(src) + b
//   kind 0: the plane index dst[0] <- src[0] along `dim`; the other two directions (in order)
//           run over lo[0]..hi[0] and lo[1]..hi[1]
//   kind 1: the z line at (dst[0], dst[1]) <- (src[0], src[1]), z over lo[0]..hi[0]
struct GhostOp {
    int kind, dim;
    int dst[2], src[2];
    int lo[2], hi[2];
    double a, b;
    long count;
};

// operations with no read-after-write dependence between them, run by one launch
constexpr int kMaxGhostOps = 8;
struct GhostWave {
    int nops;
    GhostOp ops[kMaxGhostOps];
};

void ts_deriv(const FShape &f, const Metrics &M, const double *F, double *out, int order, int d, double h,
              hipStream_t s);
// nseRHS of component c; visc multiplies the diffusion (1/Re or nu); with T != null the
// buoyancy Fb (T(k) + T(k+1)) / 2 is added to the z component
void ts_build_rhs(int c, const FShape *fv, const FShape &fp, const Metrics &M, const double *const *V,
                  const double *P, double *H, const double *h, double visc, double dt, const double *T, double Fb,
                  hipStream_t s);
// tmpRHS of the temperature (scalar_d3.cc:349-357)
void ts_build_trhs(const FShape &fp, const FShape *fv, const Metrics &M, const double *const *V, const double *T,
                   double *R, const double *h, double kappa, double dt, hipStream_t s);
void ts_nusselt(const FShape &fp, const FShape &fz, const Metrics &M, const double *Vz, const double *T,
                const double *h, double *part, double *out, hipStream_t s);
void ts_jacobi(const FShape &f, const Metrics &M, const Bulk &B, const double *V, const double *R, double *G,
               const JacobiCoef &J, hipStream_t s);
// max |residual| of the Jacobi solve over the bulk, as the bit pattern of the double in out[0];
// part holds ts_jacobi_res_parts() doubles
int ts_jacobi_res_parts();
void ts_jacobi_res(const FShape &f, const Metrics &M, const Bulk &B, const double *V, const double *R,
                   const JacobiCoef &J, double *part, unsigned long long *out, hipStream_t s);
// G = one Jacobi sweep of V (as ts_jacobi) and the max |residual| of V itself over the bulk (as
// ts_jacobi_res) in one pass; part holds ts_jacobi_vres_parts(f) doubles
long ts_jacobi_vres_parts(const FShape &f);
void ts_jacobi_vres(const FShape &f, const Metrics &M, const Bulk &B, const double *V, const double *R, double *G,
                    const JacobiCoef &J, double *part, unsigned long long *out, hipStream_t s);
void ts_ghost_wave(const FShape &f, double *F, const GhostWave &W, hipStream_t s);

// The grid factors of the Jacobi solves (hydro_d3.cc:56-61: hx2hy2 = hx^2 hy^2 ...) for spacings h,
// time step dt and Reynolds number Re (1 / Re = 0 when Re <= 0); form, vis and ms are the caller's.
JacobiCoef jacobi_coef(const double h[3], double dt, double Re);

// The bulk of a Jacobi solve (field.cc:206-219): every core entry, except that a non-periodic
// direction in which the field sits on nodes drops its two end nodes.
Bulk jacobi_bulk(const FShape &f, const int per[3]);

// The Jacobi loop of solveVx/Vy/Vz and solveT (hydro_d3.cc:370-533, scalar_d3.cc:373-594):
// sweep, the BC waves, the residual maximum over the bulk, until it is below tol.  After the first
// sweep one launch per iteration makes the next iterate in G and the residual of the current one in
// F from the same loads (ts_jacobi_vres); the next iterate gets its BCs while the host reads the
// residual through hmax and the event, and is dropped once F has converged, so F, the residuals and
// the sweep count are those of the sweep-then-check loop.  F and G are swapped as the iterates
// move; on return F holds the converged iterate, or with status 1 the last one whose residual was
// read after max_iter + 1 checks without convergence.  Returns 0 or 1, or -1 with *err set when a
// HIP call failed.  part holds max(ts_jacobi_vres_parts(f), 1) doubles; G's BC waves may still be
// queued on s on return.
struct JacobiIO {
    double *part;
    unsigned long long *dmax, *hmax;   // device maximum and its pinned host mirror
    hipEvent_t ev;
    hipStream_t s;
};
int ts_jacobi_loop(const FShape &f, const Metrics &M, const Bulk &B, double *&F, double *&G, const double *R,
                   const JacobiCoef &J, const GhostWave *bc, int nbc, const JacobiIO &io, double tol, long max_iter,
                   int *sweeps, hipError_t *err);

void ts_divergence(const FShape &fp, const FShape *fv, const Metrics &M, const double *const *V, double *out,
                   const double *h, double scale, int scaled, hipStream_t s);
void ts_project(int c, const FShape &fc, const FShape &fp, const Metrics &M, const double *Pp, double *V,
                double hc, double dt, hipStream_t s);
void ts_axpy_full(double *y, const double *x, long n, hipStream_t s);
int ts_energy_parts(const FShape &fp);
void ts_energy(const FShape &fp, const FShape *fv, const Metrics &M, const double *const *V, const double *h,
               double *part, double *out, hipStream_t s);
void ts_max(const FShape &fp, const double *F, double *part, double *out, hipStream_t s);
void ts_tgv(const FShape &f, const Metrics &M, int c, double Lx, double Ly, double Lz, double *V, hipStream_t s);
void ts_fill(double *p, long n, double v, hipStream_t s);

}  // namespace sr
