// Host-side launchers of the multigrid kernels (mg_kernels.hip).  All launches go to the given
// stream; none allocates memory or synchronises.
#pragma once

#include <hip/hip_runtime.h>

#include "mg_layout.h"

namespace sr {

template <typename T> void launch_fill_ghosts(const Lay &L, T *a, const Bc &bc, hipStream_t s);
template <typename T>
void launch_rb_sweep(const Lay &L, T *u, const T *f, const T *cx, const T *cy, const T *cz, int color,
                     const Bc &bc, bool zero_start, hipStream_t s);
template <typename T, typename TO>
// img: each computed node also writes its face and edge ghost images by *img (image rules only;
// corner ghosts are not written)
void launch_residual(const Lay &L, const T *u, const T *f, TO *t, const T *cx, const T *cy, const T *cz,
                     hipStream_t s, int color = -1, const Bc *img = nullptr);
// red0: t is zero on colour 1 and only its colour-0 entries are read (residual after a red-black
// sweep ending on colour 1, computed with color = 0)
// zero-start colour-0 half sweep of the coarse level fused into its restriction (ZS in k_restrict)
template <typename T>
struct ZeroStart {
    T *u;
    const T *cx, *cy, *cz;
    Bc bc;
};
template <typename T>
void launch_restrict(const Lay &F, const Lay &Cl, const T *t, T *fc, hipStream_t s, bool red0 = false,
                     const ZeroStart<T> *zs = nullptr);
template <typename T, typename TF>
// color >= 0: that colour's nodes only (the up leg before red-black smoothing that sweeps the other
// colour first); -1: every core node.  img: the updated nodes also write their face-ghost images by *img
// (periodic, mirror and antisymmetric rules; other rules' ghosts are left as they are)
void launch_prolong(const Lay &F, const Lay &Cl, const T *e, TF *u, int accumulate, hipStream_t s, int color = -1,
                    const Bc *img = nullptr);
template <typename T> void launch_fill(T *a, long n, T v, hipStream_t s);
template <typename T>
void launch_add_scalar(const Lay &L, T *a, const double *dv, double scale, int core, hipStream_t s);
template <typename TS, typename TD> void launch_pack(const Lay &L, const TS *ref, TD *dst, int core_only, hipStream_t s);
template <typename TS, typename TD> void launch_unpack(const Lay &L, const TS *src, TD *ref, hipStream_t s);
// every entry of the split layout: core values from the padded array, 0 elsewhere
template <typename TS, typename TD> void launch_pack_full(const Lay &L, const TS *ref, TD *dst, hipStream_t s);
int norm_parts(const Lay &L);
// out[0] = sum_core w_i w_j w_k f, out[1] = sum over every padded node of f
template <typename T>
void launch_rhs_sums(const Lay &L, const T *f, const double *wx, const double *wy, const double *wz, double *part,
                     double *out, hipStream_t s);
template <typename T> size_t coarsest_lds_bytes(const Lay &L);
// blocks of the one-node-per-thread residual norms (two partials each)
int norm_res_parts(const Lay &L);
// fterms[1] = sum f^2 (core), fterms[3] = max |f| (every padded node); the right-hand side part of
// computeError, reduced once per right-hand side
template <typename T> void launch_f_terms(const Lay &L, const T *f, double *part, double *fterms, hipStream_t s);
// q0 = weighted core sum of a reference padded array (out[0]), read in place
void launch_wsum_ref(const Lay &L, const double *ref, const double *wx, const double *wy, const double *wz,
                     double *part, double *out, hipStream_t s);
// split-layout pack of a reference padded array with the shift scale * q[0] on core nodes, and the
// right-hand side terms of computeError of the result into fterms
void launch_pack_shift(const Lay &L, const double *ref, double *dst, const double *q, double scale, double *part,
                       double *fterms, hipStream_t s);
// out[0..3] = sum |Lu-f|^2, sum f^2 (core), max |Lu-f| (core), max |f| (every padded node);
// out[1] and out[3] are copied from fterms
template <typename T>
void launch_norm_res(const Lay &L, const T *u, const T *f, const T *cx, const T *cy, const T *cz, double *part2,
                     const double *fterms, double *out, hipStream_t s, int color = -1);
// iterative-refinement step: e == nullptr computes r = f - L u_in only; otherwise u_out = u_in + e
// (core and face ghosts) and r = f - L u_out.  r is stored in FP32; out[0..3] as launch_norm_res.
void launch_outer_residual(const Lay &L, const double *u_in, const float *e, const double *f, double *u_out,
                           float *r32, const double *cx, const double *cy, const double *cz, const Bc &bc,
                           double *part2, const double *fterms, double *out, hipStream_t s);
// Krylov-accelerated mixed precision (flexible GCR(1), FP32 V-cycle preconditioner): the step
// coefficients from e = M r32 and the previous direction z (coef[0] alpha, coef[1] beta; part5 holds
// 5 doubles per block of norm_res_parts), then the step x_out = x_in + alpha (e - beta z),
// z_next = e - beta z (face-ghost images by bc and ebc); the FP64 residual of x_out is then
// launch_outer_residual's.  q1r: e ends on a colour-1 red-black half sweep of level 0 for r32, so
// L e = r32 on colour 1 and is read there instead of formed.  fb: the coefficients come from the
// two-stage final over several workgroups (k_gcr_coef_mb) instead of one workgroup
void launch_gcr_dots(const Lay &L, const float *e, const float *z, const float *r32, const float *cx,
                     const float *cy, const float *cz, double *part5, double *coef, bool q1r, hipStream_t s,
                     const FinalBufs *fb = nullptr);
void launch_gcr_step(const Lay &L, const double *x_in, const float *e, const float *z, const double *coef,
                     double *x_out, float *z_next, const Bc &bc, const Bc &ebc, hipStream_t s);
// exact coarsest solve u = P f with the dense inverse (transposed, nc x nc), ghosts filled by bc
template <typename T>
void launch_coarsest_direct(const Lay &L, T *u, const T *f, const T *pt, int nc, const Bc &bc, int *iters,
                            hipStream_t s);
template <typename T>
void launch_coarsest(const Lay &L, T *u, const T *f, const T *cx, const T *cy, const T *cz, const Bc &bc,
                     double tol, int max_count, int *iters, hipStream_t s);

// one-workgroup coarse end of the V-cycle: fills the LDS offsets of D and returns the LDS bytes
template <typename T> size_t coarse_cycle_plan(CoarseDesc &D, size_t limit);
template <typename T> int coarse_cycle_attr(size_t lds);
template <typename T>
void launch_coarse_cycle(const CoarseDesc &D, size_t lds, T *u_first, const T *f_first, int *iters, hipStream_t s);

}  // namespace sr
