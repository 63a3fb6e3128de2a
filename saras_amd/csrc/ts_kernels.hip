// gfx950 kernels of the hydro time step on the reference's padded layout (k fastest).
//
//   k_deriv        derivative::calcDerivative1_* / calcDerivative2* (lib/field/derivative.cc:121-243)
//   k_build_rhs    hydro_d3.cc:259-300 fused per velocity component: computeDiff / Re, computeNLin
//                  (vfield.cc:104-253), - grad P (sfield.cc:155-171), * dt, + V
//   k_jacobi_v     one Jacobi sweep of solveVx/Vy/Vz (hydro_d3.cc:375-391) into the second buffer
//   k_jacobi_res   the convergence residual of the same routines (hydro_d3.cc:396-410), max-reduced
//   k_jacobi_vres  the next sweep and the residual of the current iterate from the same loads
//   k_ghost_ops    a wave of ghost / wall operations (mpidata::syncData, lib/boundary/*.cc)
//   k_divergence   vfield::divergence (vfield.cc:299-320), optionally scaled (mgRHS = div / dt)
//   k_project      P += Pp and V -= dt grad Pp (hydro_d3.cc:354-362)
//   k_energy       tseries energy (tseries.cc:158-169), deterministic two-pass sum
//
// Field shapes (core sizes) follow field.h: P on nodes (N in every direction), V_c on faces in
// direction c (N-1) and nodes in the other two.  Derivative alignment: centred,
// out(r) = metric(r) (F(r+1) - F(r-1)) / (2h) over the core range along the direction.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <utility>

#include "ts_kernels.h"
#include "xcd.h"

namespace sr {

__device__ inline long fidx(const FShape &f, int i, int j, int k) {
    return ((long)(i + 1) * (f.n[1] + 2) + (j + 1)) * (f.n[2] + 2) + (k + 1);
}

// metric arrays: [dim][node?0:1][q] with q 0 coordinate, 1 xi_x, 2 xi_xx, 3 xi_x^2, index -1 at 0
__device__ inline double metv(const Metrics &M, int d, int node, int q, int i) {
    return M.m[d][node ? 0 : 1][q][i + 1];
}

// ------------------------------------------------------------------------------------------------
// derivatives (for parity tests; the fused step kernel carries its own copies)
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_deriv(FShape f, Metrics M, const double *__restrict__ F,
                                               double *__restrict__ out, int order, int d, double ih) {
    int lo[3] = {-1, -1, -1}, sz[3] = {f.n[0] + 2, f.n[1] + 2, f.n[2] + 2};
    lo[d] = 0;
    sz[d] = f.n[d];
    long total = (long)sz[0] * sz[1] * sz[2];
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int k = (int)(t % sz[2]) + lo[2];
    long r = t / sz[2];
    int j = (int)(r % sz[1]) + lo[1];
    int i = (int)(r / sz[1]) + lo[0];
    long st[3] = {(long)(f.n[1] + 2) * (f.n[2] + 2), (long)(f.n[2] + 2), 1};
    long c = fidx(f, i, j, k);
    int q = d == 0 ? i : d == 1 ? j : k;
    if (order == 1) {
        double v = (F[c + st[d]] - F[c - st[d]]) / 2.0;
        v *= ih;
        out[c] = metv(M, d, f.s[d], 1, q) * v;
    } else {
        double ih2 = ih * ih;
        double tmp = (F[c + st[d]] - F[c - st[d]]) / 2.0 * ih;
        double sec = (F[c + st[d]] - 2.0 * F[c] + F[c - st[d]]) * ih2;
        out[c] = metv(M, d, f.s[d], 2, q) * tmp + 0.5 * metv(M, d, f.s[d], 3, q) * sec;
    }
}

// ------------------------------------------------------------------------------------------------
// nseRHS of one velocity component c at its core nodes
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_build_rhs(int c, FShape fc, FShape fx, FShape fy, FShape fz, FShape fp,
                                                   Metrics M, const double *__restrict__ Vx,
                                                   const double *__restrict__ Vy, const double *__restrict__ Vz,
                                                   const double *__restrict__ P, double *__restrict__ H, double h0,
                                                   double h1, double h2, double ih0, double ih1, double ih2,
                                                   double inverseRe, double dt,
                                                   const double *__restrict__ T, double Fb) {
    // one thread per padded entry: core entries get the right-hand side, pads a copy of V (the
    // reference forms nseRHS over the whole padded array, with zero terms outside the core)
    long total = (long)(fc.n[0] + 2) * (fc.n[1] + 2) * (fc.n[2] + 2);
    const unsigned lb = xcd_block();
    long t = (long)lb * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int k = (int)(t % (fc.n[2] + 2)) - 1;
    long r = t / (fc.n[2] + 2);
    int j = (int)(r % (fc.n[1] + 2)) - 1;
    int i = (int)(r / (fc.n[1] + 2)) - 1;
    const double *V[3] = {Vx, Vy, Vz};
    const FShape *FS[3] = {&fx, &fy, &fz};
    const double *F = V[c];
    if (i < 0 || j < 0 || k < 0 || i >= fc.n[0] || j >= fc.n[1] || k >= fc.n[2]) {
        H[t] = F[t];
        return;
    }
    long st[3] = {(long)(fc.n[1] + 2) * (fc.n[2] + 2), (long)(fc.n[2] + 2), 1};
    long q0 = t;
    int idx[3] = {i, j, k};
    double hh[3] = {h0, h1, h2}, ihh[3] = {ih0, ih1, ih2};
    // computeDiff: H = d2xx + d2yy + d2zz (each added to the zero-initialised nseRHS), then / Re
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double ih = ihh[d];
        double ih2 = ih * ih;
        double fp_ = F[q0 + st[d]], fm_ = F[q0 - st[d]], f0 = F[q0];
        double tmp = (fp_ - fm_) / 2.0 * ih;
        double sec = (fp_ - 2.0 * f0 + fm_) * ih2;
        acc += metv(M, d, fc.s[d], 2, idx[d]) * tmp + 0.5 * metv(M, d, fc.s[d], 3, idx[d]) * sec;
    }
    acc *= inverseRe;
    // computeNLin: H -= (interpolated V_d) (d_d V_c) / nslices, d = x, y, z
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const FShape &fd = *FS[d];
        const double *Vd = V[d];
        double inter;
        int ns;
        if (d == c) {
            inter = Vd[fidx(fd, i, j, k)];
            ns = 1;
        } else {
            int o1[3] = {i, j, k}, o2[3] = {i, j, k}, o3[3] = {i, j, k};
            o1[d] -= 1;
            o2[c] += 1;
            o3[c] += 1;
            o3[d] -= 1;
            inter = 0.0;
            inter += Vd[fidx(fd, i, j, k)];
            inter += Vd[fidx(fd, o1[0], o1[1], o1[2])];
            inter += Vd[fidx(fd, o2[0], o2[1], o2[2])];
            inter += Vd[fidx(fd, o3[0], o3[1], o3[2])];
            ns = 4;
        }
        double ih = ihh[d];
        double dv = (F[q0 + st[d]] - F[q0 - st[d]]) / 2.0;
        dv *= ih;
        double deriv = metv(M, d, fc.s[d], 1, idx[d]) * dv;
        // / ns with ns = 1 or 4: the same as * (1 / ns), exactly
        acc -= inter * deriv * (ns == 1 ? 1.0 : 0.25);
    }
    long sp[3] = {(long)(fp.n[1] + 2) * (fp.n[2] + 2), (long)(fp.n[2] + 2), 1};
    long pq = fidx(fp, i, j, k);
    // buoyantForce::addForcing: the z face between nodes k and k + 1 gets Fb times their mean
    if (T && c == 2) {
        double inter = 0.0;
        inter += T[pq];
        inter += T[pq + 1];
        inter /= 2.0;
        acc += Fb * inter;
    }
    // - grad P on the faces of component c (sfield::gradient), * dt, + V
    double g = metv(M, c, 0, 1, idx[c]) * (P[pq + sp[c]] - P[pq]) / hh[c];
    acc -= g;
    acc *= dt;
    acc += F[q0];
    H[q0] = acc;
}

// ------------------------------------------------------------------------------------------------
// Jacobi viscous solve of component c over its bulk
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_jacobi_v(FShape f, Metrics M, Bulk B, const double *__restrict__ V,
                                                  const double *__restrict__ R, double *__restrict__ G,
                                                  JacobiCoef J) {
    // one thread per padded entry: the bulk gets the Jacobi update, every other entry 0 (the
    // reference's guess field is written on the bulk only and copied whole into V)
    unsigned n1 = (unsigned)(f.n[1] + 2), n2 = (unsigned)(f.n[2] + 2);
    unsigned total = (unsigned)(f.n[0] + 2) * n1 * n2;
    const unsigned lb = xcd_block();
    unsigned t = lb * blockDim.x + threadIdx.x;
    if (t >= total) return;
    unsigned r = t / n2;
    int k = (int)(t - r * n2) - 1;
    unsigned i0 = r / n1;
    int j = (int)(r - i0 * n1) - 1;
    int i = (int)i0 - 1;
    if (i < B.lo[0] || i > B.hi[0] || j < B.lo[1] || j > B.hi[1] || k < B.lo[2] || k > B.hi[2]) {
        G[t] = 0.0;
        return;
    }
    long st[3] = {(long)(f.n[1] + 2) * (f.n[2] + 2), (long)(f.n[2] + 2), 1};
    long q = t;
    double mx = metv(M, 0, J.ms[0], 3, i), my = metv(M, 1, J.ms[1], 3, j), mz = metv(M, 2, J.ms[2], 3, k);
    double a = J.hy2hz2 * mx * (V[q + st[0]] + V[q - st[0]]) + J.hz2hx2 * my * (V[q + st[1]] + V[q - st[1]]) +
               J.hx2hy2 * mz * (V[q + 1] + V[q - 1]);
    double d = J.hy2hz2 * mx + J.hz2hx2 * my + J.hx2hy2 * mz;
    G[q] = (a * J.rA + R[q]) / (1.0 + d * J.rD);
}

// max over the bulk of |V - 0.5 dt/Re lap V - R|: each logical block takes one contiguous range of
// the bulk (32-bit index arithmetic), so that an XCD's blocks cover neighbouring planes; one maximum
// per block into part; k_jacobi_res_final reduces them (a maximum: the same whatever the order)
__global__ void __launch_bounds__(256) k_jacobi_res(FShape f, Metrics M, Bulk B, const double *__restrict__ V,
                                                    const double *__restrict__ R, JacobiCoef J,
                                                    double *__restrict__ part) {
    unsigned sz1 = (unsigned)(B.hi[1] - B.lo[1] + 1), sz2 = (unsigned)(B.hi[2] - B.lo[2] + 1);
    unsigned total = (unsigned)(B.hi[0] - B.lo[0] + 1) * sz1 * sz2;
    long st[3] = {(long)(f.n[1] + 2) * (f.n[2] + 2), (long)(f.n[2] + 2), 1};
    const unsigned lb = xcd_block();
    // range of a block: a whole number of 256-entry strips
    const unsigned strips = (total + 255u) / 256u;
    const unsigned per = (strips + gridDim.x - 1u) / gridDim.x;
    const unsigned t0 = lb * per * 256u;
    const unsigned t1 = min(total, t0 + per * 256u);
    double e = 0.0;
    for (unsigned t = t0 + threadIdx.x; t < t1; t += 256u) {
        unsigned r = t / sz2;
        int k = (int)(t - r * sz2) + B.lo[2];
        unsigned i0 = r / sz1;
        int j = (int)(r - i0 * sz1) + B.lo[1];
        int i = (int)i0 + B.lo[0];
        long q = fidx(f, i, j, k);
        double mx = metv(M, 0, J.ms[0], 3, i), my = metv(M, 1, J.ms[1], 3, j), mz = metv(M, 2, J.ms[2], 3, k);
        double lap = V[q] - (mx * (V[q + st[0]] - 2.0 * V[q] + V[q - st[0]]) * J.ih2[0] +
                             my * (V[q + st[1]] - 2.0 * V[q] + V[q - st[1]]) * J.ih2[1] +
                             mz * (V[q + 1] - 2.0 * V[q] + V[q - 1]) * J.ih2[2]) * J.lapS;
        e = fmax(e, fabs(lap - R[q]));
    }
    __shared__ double s[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) e = fmax(e, __shfl_xor(e, o, 64));
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) part[lb] = fmax(fmax(s[0], s[1]), fmax(s[2], s[3]));
}

// One Jacobi sweep G = J(V), as k_jacobi_v, and from the same loads the convergence residual of V
// itself, as k_jacobi_res: the sweep that follows an iterate and that iterate's residual read the
// same neighbours.  Both expressions are those of the two kernels, so the iterates and the residuals
// are theirs.  A logical block covers 1024 consecutive padded entries (four per thread) and stores
// the maximum residual of its bulk entries in part.
__global__ void __launch_bounds__(256) k_jacobi_vres(FShape f, Metrics M, Bulk B, const double *__restrict__ V,
                                                     const double *__restrict__ R, double *__restrict__ G,
                                                     JacobiCoef J, double *__restrict__ part) {
    const unsigned n1 = (unsigned)(f.n[1] + 2), n2 = (unsigned)(f.n[2] + 2);
    const unsigned total = (unsigned)(f.n[0] + 2) * n1 * n2;
    const unsigned lb = xcd_block();
    const long st[3] = {(long)(f.n[1] + 2) * (f.n[2] + 2), (long)(f.n[2] + 2), 1};
    double e = 0.0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const unsigned t = lb * 1024u + (unsigned)u * 256u + threadIdx.x;
        if (t >= total) break;
        unsigned r = t / n2;
        int k = (int)(t - r * n2) - 1;
        unsigned i0 = r / n1;
        int j = (int)(r - i0 * n1) - 1;
        int i = (int)i0 - 1;
        if (i < B.lo[0] || i > B.hi[0] || j < B.lo[1] || j > B.hi[1] || k < B.lo[2] || k > B.hi[2]) {
            G[t] = 0.0;
            continue;
        }
        const long q = t;
        const double vxp = V[q + st[0]], vxm = V[q - st[0]], vyp = V[q + st[1]], vym = V[q - st[1]];
        const double vzp = V[q + 1], vzm = V[q - 1], v0 = V[q], rq = R[q];
        double mx = metv(M, 0, J.ms[0], 3, i), my = metv(M, 1, J.ms[1], 3, j), mz = metv(M, 2, J.ms[2], 3, k);
        double a = J.hy2hz2 * mx * (vxp + vxm) + J.hz2hx2 * my * (vyp + vym) + J.hx2hy2 * mz * (vzp + vzm);
        double d = J.hy2hz2 * mx + J.hz2hx2 * my + J.hx2hy2 * mz;
        G[q] = (a * J.rA + rq) / (1.0 + d * J.rD);
        double lap = v0 - (mx * (vxp - 2.0 * v0 + vxm) * J.ih2[0] + my * (vyp - 2.0 * v0 + vym) * J.ih2[1] +
                           mz * (vzp - 2.0 * v0 + vzm) * J.ih2[2]) * J.lapS;
        e = fmax(e, fabs(lap - rq));
    }
    __shared__ double sm[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) e = fmax(e, __shfl_xor(e, o, 64));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) part[lb] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
}

// maximum of the block maxima, stored as the bit pattern of the double; one block of up to 1024
// threads (a maximum: the same whatever the order)
__global__ void __launch_bounds__(1024) k_jacobi_res_final(const double *__restrict__ part, int n,
                                                           unsigned long long *__restrict__ out) {
    double m = 0.0;
#pragma unroll 8
    for (int q = threadIdx.x; q < n; q += blockDim.x) m = fmax(m, part[q]);
    __shared__ double s[16];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = s[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = fmax(r, s[w]);
        out[0] = (unsigned long long)__double_as_longlong(r);
    }
}

// ------------------------------------------------------------------------------------------------
// ghost / wall operations: dst = a * F(src) + b over a plane (dim fixed) or a z-line (x, y fixed)
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_ghost_ops(FShape f, double *__restrict__ F, GhostWave W) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int o = 0;
    while (o < W.nops && t >= W.ops[o].count) {
        t -= W.ops[o].count;
        o++;
    }
    if (o >= W.nops) return;
    const GhostOp &op = W.ops[o];
    int p[3], s[3];
    if (op.kind == 0) {      // plane: fixed index along op.dim, ranges on the other two
        int a = op.dim == 0 ? 1 : 0, b = op.dim == 2 ? 1 : 2;
        int na = op.hi[0] - op.lo[0] + 1;
        int ia = op.lo[0] + (int)(t % na), ib = op.lo[1] + (int)(t / na);
        p[op.dim] = op.dst[0];
        s[op.dim] = op.src[0];
        p[a] = s[a] = ia;
        p[b] = s[b] = ib;
    } else {                 // z-line at fixed (x, y)
        int kk = op.lo[0] + (int)t;
        p[0] = op.dst[0]; p[1] = op.dst[1]; p[2] = kk;
        s[0] = op.src[0]; s[1] = op.src[1]; s[2] = kk;
    }
    double v = op.a == 0.0 ? 0.0 : op.a * F[fidx(f, s[0], s[1], s[2])];
    F[fidx(f, p[0], p[1], p[2])] = v + op.b;
}

// ------------------------------------------------------------------------------------------------
// temperature right-hand side: core nodes T + dt (kappa sum d2 T - sum_d Vbar_d dT/dx_d), pads T
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_build_trhs(FShape fp, FShape fx, FShape fy, FShape fz, Metrics M,
                                                    const double *__restrict__ Vx, const double *__restrict__ Vy,
                                                    const double *__restrict__ Vz, const double *__restrict__ T,
                                                    double *__restrict__ R, double ih0, double ih1, double ih2,
                                                    double kappa, double dt) {
    long total = (long)(fp.n[0] + 2) * (fp.n[1] + 2) * (fp.n[2] + 2);
    const unsigned lb = xcd_block();
    long t = (long)lb * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int k = (int)(t % (fp.n[2] + 2)) - 1;
    long r = t / (fp.n[2] + 2);
    int j = (int)(r % (fp.n[1] + 2)) - 1;
    int i = (int)(r / (fp.n[1] + 2)) - 1;
    if (i < 0 || j < 0 || k < 0 || i >= fp.n[0] || j >= fp.n[1] || k >= fp.n[2]) {
        R[t] = T[t];
        return;
    }
    long st[3] = {(long)(fp.n[1] + 2) * (fp.n[2] + 2), (long)(fp.n[2] + 2), 1};
    int idx[3] = {i, j, k};
    double ihh[3] = {ih0, ih1, ih2};
    const double *V[3] = {Vx, Vy, Vz};
    const FShape *FS[3] = {&fx, &fy, &fz};
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double ih = ihh[d];
        double ih2 = ih * ih;
        double fpv = T[t + st[d]], fmv = T[t - st[d]], f0 = T[t];
        double tmp = (fpv - fmv) / 2.0 * ih;
        double sec = (fpv - 2.0 * f0 + fmv) * ih2;
        acc += metv(M, d, 1, 2, idx[d]) * tmp + 0.5 * metv(M, d, 1, 3, idx[d]) * sec;
    }
    acc *= kappa;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int o[3] = {i, j, k};
        o[d] -= 1;
        double inter = 0.0;
        inter += V[d][fidx(*FS[d], i, j, k)];
        inter += V[d][fidx(*FS[d], o[0], o[1], o[2])];
        double ih = ihh[d];
        double dv = (T[t + st[d]] - T[t - st[d]]) / 2.0;
        dv *= ih;
        double deriv = metv(M, d, 1, 1, idx[d]) * dv;
        acc -= inter * deriv / 2.0;
    }
    acc *= dt;
    acc += T[t];
    R[t] = acc;
}

// Nusselt partials: per block sums of (volume) and (w T volume), w the mean of the two z faces
__global__ void __launch_bounds__(256) k_nusselt(FShape fp, FShape fz, Metrics M, const double *__restrict__ Vz,
                                                 const double *__restrict__ T, double h0, double h1, double h2,
                                                 double *__restrict__ part) {
    long total = (long)fp.n[0] * fp.n[1] * fp.n[2];
    const unsigned lb = xcd_block();
    long t = (long)lb * blockDim.x + threadIdx.x;
    double vol = 0.0, uzt = 0.0;
    if (t < total) {
        int k = (int)(t % fp.n[2]);
        long r = t / fp.n[2];
        int j = (int)(r % fp.n[1]);
        int i = (int)(r / fp.n[1]);
        double dv = (h0 / metv(M, 0, 0, 1, i)) * (h1 / metv(M, 1, 0, 1, j)) * (h2 / metv(M, 2, 0, 1, k));
        double w = (Vz[fidx(fz, i, j, k)] + Vz[fidx(fz, i, j, k - 1)]) / 2.0;
        vol = dv;
        uzt = w * T[fidx(fp, i, j, k)] * dv;
    }
    __shared__ double sv[4], su[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        vol += __shfl_xor(vol, o, 64);
        uzt += __shfl_xor(uzt, o, 64);
    }
    if ((threadIdx.x & 63) == 0) { sv[threadIdx.x >> 6] = vol; su[threadIdx.x >> 6] = uzt; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0;
        for (int w = 0; w < 4; w++) { a += sv[w]; b += su[w]; }
        part[2 * (long)lb] = a;
        part[2 * (long)lb + 1] = b;
    }
}

// ------------------------------------------------------------------------------------------------
// divergence over the node core (pads untouched), times `scale`
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_divergence(FShape fp, FShape fx, FShape fy, FShape fz, Metrics M,
                                                    const double *__restrict__ Vx, const double *__restrict__ Vy,
                                                    const double *__restrict__ Vz, double *__restrict__ out,
                                                    double h0, double h1, double h2, double scale, int scaled) {
    long total = (long)fp.n[0] * fp.n[1] * fp.n[2];
    const unsigned lb = xcd_block();
    long t = (long)lb * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int k = (int)(t % fp.n[2]);
    long r = t / fp.n[2];
    int j = (int)(r % fp.n[1]);
    int i = (int)(r / fp.n[1]);
    double a = metv(M, 0, 1, 1, i) * (Vx[fidx(fx, i, j, k)] - Vx[fidx(fx, i - 1, j, k)]) / h0;
    double b = metv(M, 1, 1, 1, j) * (Vy[fidx(fy, i, j, k)] - Vy[fidx(fy, i, j - 1, k)]) / h1;
    double c = metv(M, 2, 1, 1, k) * (Vz[fidx(fz, i, j, k)] - Vz[fidx(fz, i, j, k - 1)]) / h2;
    double d = a + b + c;
    out[fidx(fp, i, j, k)] = scaled ? d * scale : d;
}

// ------------------------------------------------------------------------------------------------
// projection: V_c(core) -= dt * grad_c Pp  (pads of the gradient are zero, so pads of V are kept)
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_project(int c, FShape fc, FShape fp, Metrics M, const double *__restrict__ Pp,
                                                 double *__restrict__ V, double hc, double dt) {
    long total = (long)fc.n[0] * fc.n[1] * fc.n[2];
    const unsigned lb = xcd_block();
    long t = (long)lb * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int k = (int)(t % fc.n[2]);
    long r = t / fc.n[2];
    int j = (int)(r % fc.n[1]);
    int i = (int)(r / fc.n[1]);
    int idx[3] = {i, j, k};
    long sp[3] = {(long)(fp.n[1] + 2) * (fp.n[2] + 2), (long)(fp.n[2] + 2), 1};
    long pq = fidx(fp, i, j, k);
    double g = metv(M, c, 0, 1, idx[c]) * (Pp[pq + sp[c]] - Pp[pq]) / hc;
    g *= dt;
    long q = fidx(fc, i, j, k);
    V[q] = V[q] - g;
}

__global__ void __launch_bounds__(256) k_fill_d(double *__restrict__ p, long n, double v) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void __launch_bounds__(256) k_axpy_full(double *__restrict__ y, const double *__restrict__ x, long n) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] += x[t];
}

// ------------------------------------------------------------------------------------------------
// energy partials: per block sums of (volume) and (energy density * volume), fixed order
// ------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_energy(FShape fp, FShape fx, FShape fy, FShape fz, Metrics M,
                                                const double *__restrict__ Vx, const double *__restrict__ Vy,
                                                const double *__restrict__ Vz, double h0, double h1, double h2,
                                                double *__restrict__ part) {
    long total = (long)fp.n[0] * fp.n[1] * fp.n[2];
    const unsigned lb = xcd_block();
    long t = (long)lb * blockDim.x + threadIdx.x;
    double vol = 0.0, en = 0.0;
    if (t < total) {
        int k = (int)(t % fp.n[2]);
        long r = t / fp.n[2];
        int j = (int)(r % fp.n[1]);
        int i = (int)(r / fp.n[1]);
        double dv = (h0 / metv(M, 0, 0, 1, i)) * (h1 / metv(M, 1, 0, 1, j)) * (h2 / metv(M, 2, 0, 1, k));
        double u = (Vx[fidx(fx, i - 1, j, k)] + Vx[fidx(fx, i, j, k)]) / 2.0;
        double v = (Vy[fidx(fy, i, j - 1, k)] + Vy[fidx(fy, i, j, k)]) / 2.0;
        double w = (Vz[fidx(fz, i, j, k - 1)] + Vz[fidx(fz, i, j, k)]) / 2.0;
        vol = dv;
        en = (u * u + v * v + w * w) * dv;
    }
    __shared__ double sv[4], se[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        vol += __shfl_xor(vol, o, 64);
        en += __shfl_xor(en, o, 64);
    }
    if ((threadIdx.x & 63) == 0) { sv[threadIdx.x >> 6] = vol; se[threadIdx.x >> 6] = en; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0;
        for (int w = 0; w < 4; w++) { a += sv[w]; b += se[w]; }
        part[2 * (long)lb] = a;
        part[2 * (long)lb + 1] = b;
    }
}

__global__ void __launch_bounds__(256) k_energy_final(const double *__restrict__ part, int n, double *__restrict__ out) {
    double a = 0, b = 0;
#pragma unroll 8
    for (int q = threadIdx.x; q < n; q += blockDim.x) {
        a += part[2 * (long)q];
        b += part[2 * (long)q + 1];
    }
    __shared__ double sa[4], sb[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    if ((threadIdx.x & 63) == 0) { sa[threadIdx.x >> 6] = a; sb[threadIdx.x >> 6] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0, y = 0;
        for (int w = 0; w < 4; w++) { x += sa[w]; y += sb[w]; }
        out[0] = x;
        out[1] = y;
    }
}

// max over the node core of a field (fxMax); the values are arbitrary signed doubles, reduced in
// a fixed order per block then by one block
__global__ void __launch_bounds__(256) k_max_partial(FShape fp, const double *__restrict__ F, double *__restrict__ part) {
    long total = (long)fp.n[0] * fp.n[1] * fp.n[2];
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    double m = -1.0e300;
    if (t < total) {
        int k = (int)(t % fp.n[2]);
        long r = t / fp.n[2];
        int j = (int)(r % fp.n[1]);
        int i = (int)(r / fp.n[1]);
        m = F[fidx(fp, i, j, k)];
    }
    __shared__ double s[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(s[0], s[1]), fmax(s[2], s[3]));
}

__global__ void __launch_bounds__(256) k_max_final(const double *__restrict__ part, int n, double *__restrict__ out) {
    double m = -1.0e300;
#pragma unroll 8
    for (int q = threadIdx.x; q < n; q += blockDim.x) m = fmax(m, part[q]);
    __shared__ double s[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = fmax(fmax(s[0], s[1]), fmax(s[2], s[3]));
}

__global__ void k_tgv(FShape f, Metrics M, int c, double Lx, double Ly, double Lz, double *__restrict__ V) {
    long total = (long)(f.n[0] + 2) * (f.n[1] + 2) * (f.n[2] + 2);
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int k = (int)(t % (f.n[2] + 2)) - 1;
    long r = t / (f.n[2] + 2);
    int j = (int)(r % (f.n[1] + 2)) - 1;
    int i = (int)(r / (f.n[1] + 2)) - 1;
    double x = metv(M, 0, f.s[0], 0, i), y = metv(M, 1, f.s[1], 0, j), z = metv(M, 2, f.s[2], 0, k);
    const double pi = 3.14159265358979323846;
    double v = 0.0;
    if (c == 0) v = sin(2.0 * pi * x / Lx) * cos(2.0 * pi * y / Ly) * cos(2.0 * pi * z / Lz);
    if (c == 1) v = -cos(2.0 * pi * x / Lx) * sin(2.0 * pi * y / Ly) * cos(2.0 * pi * z / Lz);
    V[t] = v;
}

// ------------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------------
static inline unsigned nb(long n) { return (unsigned)((n + 255) / 256); }

void ts_deriv(const FShape &f, const Metrics &M, const double *F, double *out, int order, int d, double h,
              hipStream_t s) {
    long sz[3] = {f.n[0] + 2L, f.n[1] + 2L, f.n[2] + 2L};
    sz[d] = f.n[d];
    hipLaunchKernelGGL(k_deriv, dim3(nb(sz[0] * sz[1] * sz[2])), dim3(256), 0, s, f, M, F, out, order, d, 1.0 / h);
}

void ts_build_rhs(int c, const FShape *fv, const FShape &fp, const Metrics &M, const double *const *V,
                  const double *P, double *H, const double *h, double visc, double dt, const double *T, double Fb,
                  hipStream_t s) {
    const FShape &fc = fv[c];
    long n = (long)(fc.n[0] + 2) * (fc.n[1] + 2) * (fc.n[2] + 2);
    // the inverse grid spacings are formed on the host, like derivative's ihx, ihy, ihz (derivative.cc:66-68)
    hipLaunchKernelGGL(k_build_rhs, dim3(nb(n)), dim3(256), 0, s, c, fc, fv[0], fv[1], fv[2], fp, M, V[0], V[1],
                       V[2], P, H, h[0], h[1], h[2], 1.0 / h[0], 1.0 / h[1], 1.0 / h[2], visc, dt, T, Fb);
}

void ts_build_trhs(const FShape &fp, const FShape *fv, const Metrics &M, const double *const *V, const double *T,
                   double *R, const double *h, double kappa, double dt, hipStream_t s) {
    long n = (long)(fp.n[0] + 2) * (fp.n[1] + 2) * (fp.n[2] + 2);
    hipLaunchKernelGGL(k_build_trhs, dim3(nb(n)), dim3(256), 0, s, fp, fv[0], fv[1], fv[2], M, V[0], V[1], V[2], T, R,
                       1.0 / h[0], 1.0 / h[1], 1.0 / h[2], kappa, dt);
}

void ts_nusselt(const FShape &fp, const FShape &fz, const Metrics &M, const double *Vz, const double *T,
                const double *h, double *part, double *out, hipStream_t s) {
    int n = ts_energy_parts(fp);
    hipLaunchKernelGGL(k_nusselt, dim3(n), dim3(256), 0, s, fp, fz, M, Vz, T, h[0], h[1], h[2], part);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, s, part, n, out);
}

// the grid-uniform quotients of hydro_d3.cc:375-410 (form 0: 1/Re; form 1: the diffusivity vis,
// scalar_d3.cc:373-531), formed once on the host
static JacobiCoef with_quotients(JacobiCoef J) {
    if (J.form == 0) {
        J.rA = J.dt / (J.hx2hy2hz2 * 2.0 * J.Re);
        J.rD = J.dt / (J.Re * J.hx2hy2hz2);
    } else {
        J.rA = J.dt * J.vis / (J.hx2hy2hz2 * 2.0);
        J.rD = J.dt * J.vis / J.hx2hy2hz2;
    }
    J.ih2[0] = 1.0 / (J.hx * J.hx);
    J.ih2[1] = 1.0 / (J.hy * J.hy);
    J.ih2[2] = 1.0 / (J.hz * J.hz);
    J.lapS = 0.5 * J.dt * (J.form == 0 ? J.inverseRe : J.vis);
    return J;
}

void ts_jacobi(const FShape &f, const Metrics &M, const Bulk &B, const double *V, const double *R, double *G,
               const JacobiCoef &J, hipStream_t s) {
    long n = (long)(f.n[0] + 2) * (f.n[1] + 2) * (f.n[2] + 2);
    hipLaunchKernelGGL(k_jacobi_v, dim3(nb(n)), dim3(256), 0, s, f, M, B, V, R, G, with_quotients(J));
}

int ts_jacobi_res_parts() { return 1024; }

void ts_jacobi_res(const FShape &f, const Metrics &M, const Bulk &B, const double *V, const double *R,
                   const JacobiCoef &J, double *part, unsigned long long *out, hipStream_t s) {
    long n = (long)(B.hi[0] - B.lo[0] + 1) * (B.hi[1] - B.lo[1] + 1) * (B.hi[2] - B.lo[2] + 1);
    unsigned g = nb(n);
    if (g > (unsigned)ts_jacobi_res_parts()) g = (unsigned)ts_jacobi_res_parts();
    if (g == 0) g = 1;
    hipLaunchKernelGGL(k_jacobi_res, dim3(g), dim3(256), 0, s, f, M, B, V, R, with_quotients(J), part);
    hipLaunchKernelGGL(k_jacobi_res_final, dim3(1), dim3(1024), 0, s, part, (int)g, out);
}

long ts_jacobi_vres_parts(const FShape &f) {
    return ((long)(f.n[0] + 2) * (f.n[1] + 2) * (f.n[2] + 2) + 1023) / 1024;
}

void ts_jacobi_vres(const FShape &f, const Metrics &M, const Bulk &B, const double *V, const double *R, double *G,
                    const JacobiCoef &J, double *part, unsigned long long *out, hipStream_t s) {
    unsigned g = (unsigned)ts_jacobi_vres_parts(f);
    hipLaunchKernelGGL(k_jacobi_vres, dim3(g), dim3(256), 0, s, f, M, B, V, R, G, with_quotients(J), part);
    hipLaunchKernelGGL(k_jacobi_res_final, dim3(1), dim3(1024), 0, s, part, (int)g, out);
}

void ts_ghost_wave(const FShape &f, double *F, const GhostWave &W, hipStream_t s) {
    long n = 0;
    for (int o = 0; o < W.nops; o++) n += W.ops[o].count;
    if (n > 0) hipLaunchKernelGGL(k_ghost_ops, dim3(nb(n)), dim3(256), 0, s, f, F, W);
}

JacobiCoef jacobi_coef(const double h[3], double dt, double Re) {
    JacobiCoef J{};
    J.hx = h[0];
    J.hy = h[1];
    J.hz = h[2];
    J.hx2hy2 = std::pow(J.hx, 2.0) * std::pow(J.hy, 2.0);
    J.hy2hz2 = std::pow(J.hy, 2.0) * std::pow(J.hz, 2.0);
    J.hz2hx2 = std::pow(J.hz, 2.0) * std::pow(J.hx, 2.0);
    J.hx2hy2hz2 = std::pow(J.hx, 2.0) * std::pow(J.hy, 2.0) * std::pow(J.hz, 2.0);
    J.dt = dt;
    J.Re = Re;
    J.inverseRe = Re > 0.0 ? 1.0 / Re : 0.0;
    return J;
}

Bulk jacobi_bulk(const FShape &f, const int per[3]) {
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

int ts_jacobi_loop(const FShape &f, const Metrics &M, const Bulk &B, double *&F, double *&G, const double *R,
                   const JacobiCoef &J, const GhostWave *bc, int nbc, const JacobiIO &io, double tol, long max_iter,
                   int *sweeps, hipError_t *err) {
#define JCHK(x)                      \
    do {                             \
        hipError_t e_ = (x);         \
        if (e_ != hipSuccess) {      \
            *err = e_;               \
            return -1;               \
        }                            \
    } while (0)
    long iterCount = 0;
    int n = 0;
    ts_jacobi(f, M, B, F, R, G, J, io.s);
    std::swap(F, G);
    for (int w = 0; w < nbc; w++) ts_ghost_wave(f, F, bc[w], io.s);
    int status = 0;
    while (true) {
        ts_jacobi_vres(f, M, B, F, R, G, J, io.part, io.dmax, io.s);
        JCHK(hipMemcpyAsync(io.hmax, io.dmax, sizeof(unsigned long long), hipMemcpyDeviceToHost, io.s));
        JCHK(hipEventRecord(io.ev, io.s));
        for (int w = 0; w < nbc; w++) ts_ghost_wave(f, G, bc[w], io.s);
        JCHK(hipGetLastError());
        JCHK(hipEventSynchronize(io.ev));
        double maxError;
        unsigned long long bits = *io.hmax;
        std::memcpy(&maxError, &bits, sizeof(double));
        n += 1;
        if (maxError < tol) break;
        iterCount += 1;
        if (iterCount > max_iter) {
            status = 1;
            break;
        }
        std::swap(F, G);
    }
#undef JCHK
    *sweeps = n;
    return status;
}

void ts_divergence(const FShape &fp, const FShape *fv, const Metrics &M, const double *const *V, double *out,
                   const double *h, double scale, int scaled, hipStream_t s) {
    long n = (long)fp.n[0] * fp.n[1] * fp.n[2];
    hipLaunchKernelGGL(k_divergence, dim3(nb(n)), dim3(256), 0, s, fp, fv[0], fv[1], fv[2], M, V[0], V[1], V[2],
                       out, h[0], h[1], h[2], scale, scaled);
}

void ts_project(int c, const FShape &fc, const FShape &fp, const Metrics &M, const double *Pp, double *V,
                double hc, double dt, hipStream_t s) {
    long n = (long)fc.n[0] * fc.n[1] * fc.n[2];
    hipLaunchKernelGGL(k_project, dim3(nb(n)), dim3(256), 0, s, c, fc, fp, M, Pp, V, hc, dt);
}

void ts_axpy_full(double *y, const double *x, long n, hipStream_t s) {
    hipLaunchKernelGGL(k_axpy_full, dim3(nb(n)), dim3(256), 0, s, y, x, n);
}


void ts_fill(double *p, long n, double v, hipStream_t s) {
    if (n > 0) hipLaunchKernelGGL(k_fill_d, dim3(nb(n)), dim3(256), 0, s, p, n, v);
}

int ts_energy_parts(const FShape &fp) { return (int)nb((long)fp.n[0] * fp.n[1] * fp.n[2]); }

void ts_energy(const FShape &fp, const FShape *fv, const Metrics &M, const double *const *V, const double *h,
               double *part, double *out, hipStream_t s) {
    int n = ts_energy_parts(fp);
    hipLaunchKernelGGL(k_energy, dim3(n), dim3(256), 0, s, fp, fv[0], fv[1], fv[2], M, V[0], V[1], V[2], h[0],
                       h[1], h[2], part);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(256), 0, s, part, n, out);
}

void ts_max(const FShape &fp, const double *F, double *part, double *out, hipStream_t s) {
    int n = ts_energy_parts(fp);
    hipLaunchKernelGGL(k_max_partial, dim3(n), dim3(256), 0, s, fp, F, part);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, s, part, n, out);
}

void ts_tgv(const FShape &f, const Metrics &M, int c, double Lx, double Ly, double Lz, double *V, hipStream_t s) {
    long n = (long)(f.n[0] + 2) * (f.n[1] + 2) * (f.n[2] + 2);
    hipLaunchKernelGGL(k_tgv, dim3(nb(n)), dim3(256), 0, s, f, M, c, Lx, Ly, Lz, V);
}

}  // namespace sr
