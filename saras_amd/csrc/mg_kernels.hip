// gfx950 kernels of the multigrid V-cycle on the red-black split layout (mg_layout.h).
//
// Each kernel restates one operator of lib/poisson/poisson_d3.cc on one level:
//   k_rb_sweep       smooth() Gauss-Seidel sweep (poisson_d3.cc:137-150), red-black order, with the
//                    face-ghost images of every updated node written by the updating thread
//   k_fill_ghosts    imposeBC() / updatePads() (poisson_d3.cc:575-686), composed per ghost node
//   k_residual       computeResidual() (poisson_d3.cc:90-111)
//   k_restrict       coarsen(): 27-point full weighting (poisson_d3.cc:250-298)
//   k_prolong        prolong() (poisson_d3.cc:301-388), optionally fused with += smoothedPres
//   k_norm_*         computeError() (poisson_d3.cc:391-459): deterministic two-pass reductions
//   k_coarsest       solve() (poisson_d3.cc:190-247): the whole coarsest-level iteration in LDS
//
// The discrete operator is the reference's, rewritten with per-index coefficients
//   L u = sum_d lp_d u(+1) + lm_d u(-1) - lc_d u(0),
//   lp = xix2/h^2 + xixx/(2h),  lm = xix2/h^2 - xixx/(2h),  lc = 2 xix2/h^2,
// and the Gauss-Seidel update u = (sum_d lp_d u(+1) + lm_d u(-1) - f) / sum_d lc_d, which equals the
// reference quotient of poisson_d3.cc:140-147 after dividing numerator and denominator by 2 hx^2hy^2hz^2.
#include "mg_kernels.h"
#include "mg_state.h"
#include "xcd.h"

#include <hip/hip_runtime.h>

namespace sr {

// Launch geometry over whole planes: logical block b covers slots [256 (b % bpp), +256) of plane
// ip0 + b / bpp.  The plane index is wave-uniform (scalar division); the split of the in-plane
// slot index uses a float reciprocal with one correction step, exact for r < 2^24.
struct Geo {
    unsigned bpp, per_plane, ip0, nplanes, total_blocks;
    unsigned d1;
    float inv1;
};

__device__ inline unsigned fdivu(unsigned r, unsigned d, float inv) {
    unsigned q = (unsigned)((float)r * inv);
    int rem = (int)r - (int)(q * d);
    q += (unsigned)(rem >= (int)d) - (unsigned)(rem < 0);
    return q;
}

// Grid-stride order for reductions over a whole level: the blocks of XCD x (blockIdx % 8) walk the
// x-th contiguous eighth of the logical blocks, so the neighbour planes a block reads stay in that
// XCD's L2.  Returns the first logical block, the stride and the end of this XCD's range.
__device__ inline void xcd_range(unsigned total, unsigned &first, unsigned &stride, unsigned &end) {
    unsigned x = blockIdx.x & 7u, w = blockIdx.x >> 3, per = gridDim.x >> 3;
    unsigned chunk = (total + 7u) / 8u;
    first = x * chunk + w;
    stride = per;
    end = min(total, (x + 1u) * chunk);
}

// t = (a * n1 + b) * n2 + c for t < 2^24, split with float-reciprocal divisions (fdivu) instead of
// integer divisions: the single-workgroup kernels are bound by their VALU issue rate
struct Div3 {
    unsigned n1, n2;
    float i1, i2;
};
__device__ inline Div3 div3(int n1, int n2) {
    return Div3{(unsigned)n1, (unsigned)n2, 1.0f / (float)n1, 1.0f / (float)n2};
}
__device__ inline void split3(const Div3 &D, int t, int &a, int &b, int &c) {
    unsigned r = fdivu((unsigned)t, D.n2, D.i2);
    c = t - (int)(r * D.n2);
    unsigned q = fdivu(r, D.n1, D.i1);
    b = (int)(r - q * D.n1);
    a = (int)q;
}

__device__ inline bool geo_slot(const Geo &g, unsigned b, unsigned &ip, unsigned &r) {
    unsigned p = b / g.bpp;
    r = (b - p * g.bpp) * 256u + threadIdx.x;
    ip = g.ip0 + p;
    return p < g.nplanes && r < g.per_plane;
}

// ------------------------------------------------------------------------------------------------
// ghost composition (imposeBC order x, y, z; z outermost)
// ------------------------------------------------------------------------------------------------
// Where ghost (ip, jp, kp) takes its value in the composition: 0 it is not written, 1 it is
// sign * a[src] (src a core node), 2 it is the constant cval.
template <typename T>
__device__ inline int ghost_source(const Lay &L, int ip, int jp, int kp, const Bc &bc, long &src, T &sign, T &cval) {
    int n[3] = {L.nx, L.ny, L.nz};
    int q[3] = {ip, jp, kp};
    T sg = T(1);
    bool moved = false;
    for (int d = 2; d >= 0; --d) {
        int g = q[d];
        if (g != 0 && g != n[d] + 1) continue;
        int r = bc.rule[d];
        if (r == G_KEEP) continue;
        if (r == G_ZERO) { cval = T(0); return 2; }
        if (r == G_WALL) {
            const double *w = bc.wall[2 * d + (g == 0 ? 0 : 1)];
            T v = T(0);
            if (w) {
                long wi = d == 0 ? (long)q[1] * (L.nz + 2) + q[2]
                        : d == 1 ? (long)q[0] * (L.nz + 2) + q[2]
                                 : (long)q[0] * (L.ny + 2) + q[1];
                v = (T)w[wi];
            }
            cval = sg * v;
            return 2;
        }
        q[d] = ghost_src(r, g, n[d]);
        if (r == G_ANTI) sg = -sg;
        moved = true;
    }
    if (!moved) return 0;
    src = lay_idx(L, q[0], q[1], q[2]);
    sign = sg;
    return 1;
}

template <typename T>
__device__ inline bool ghost_resolve(const Lay &L, const T *a, int ip, int jp, int kp, const Bc &bc, T &out) {
    long src = 0;
    T sign = T(1), cval = T(0);
    int kind = ghost_source<T>(L, ip, jp, kp, bc, src, sign, cval);
    if (kind == 0) return false;
    out = kind == 1 ? sign * a[src] : cval;
    return true;
}

// One thread per ghost node: z faces (full ip, jp), then y faces (full ip, core kp), then x faces.
template <typename T>
__global__ void __launch_bounds__(256) k_fill_ghosts(Lay L, T *__restrict__ a, Bc bc, long total) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int NX = L.nx + 2, NY = L.ny + 2;
    long nzf = 2L * NX * NY;
    long nyf = 2L * NX * L.nz;
    int ip, jp, kp;
    if (t < nzf) {
        long r = t;
        kp = (r & 1) ? L.nz + 1 : 0;
        r >>= 1;
        jp = (int)(r % NY);
        ip = (int)(r / NY);
    } else if (t < nzf + nyf) {
        long r = t - nzf;
        kp = 1 + (int)(r % L.nz);
        r /= L.nz;
        jp = (r & 1) ? L.ny + 1 : 0;
        ip = (int)(r >> 1);
    } else {
        long r = t - nzf - nyf;
        kp = 1 + (int)(r % L.nz);
        r /= L.nz;
        jp = 1 + (int)(r % L.ny);
        r /= L.ny;
        ip = r ? L.nx + 1 : 0;
    }
    T v;
    if (ghost_resolve(L, a, ip, jp, kp, bc, v)) a[lay_idx(L, ip, jp, kp)] = v;
}

// ------------------------------------------------------------------------------------------------
// red-black Gauss-Seidel half sweep
// ------------------------------------------------------------------------------------------------
// Face-ghost images of an updated node: a node whose padded index is the source of a ghost of a
// periodic / mirror / anti direction writes that ghost (same colour, so no other thread of this
// half sweep reads it).  Edge and corner ghosts are not read by the 7-point stencil.
// Diagonal of the 7-point operator at a node from the three directions' coefficient quadruples
// (c+, c-, c0, pad).  The FP64 quadruples store c0 = c+ + c- (rounded once, at set-up), so the sum
// is formed from c+ and c-, which the stencil loads anyway, with the same value as from c0; the FP32
// quadruples hold each value rounded separately and use c0.
template <typename T>
__device__ __forceinline__ T diag3(const T *X, const T *Y, const T *Z) {
    if constexpr (sizeof(T) == 8) return (X[0] + X[1]) + (Y[0] + Y[1]) + (Z[0] + Z[1]);
    else return X[2] + Y[2] + Z[2];
}

template <typename T>
__device__ inline void write_images(const Lay &L, T *u, int ip, int jp, int kp, T v, const Bc &bc) {
    int n[3] = {L.nx, L.ny, L.nz};
    int p[3] = {ip, jp, kp};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        int r = bc.rule[d];
        if (r > G_ANTI) continue;
        T w = (r == G_ANTI) ? -v : v;
        int g0 = ghost_src(r, 0, n[d]), g1 = ghost_src(r, n[d] + 1, n[d]);
        if (p[d] == g0) {
            int q[3] = {p[0], p[1], p[2]};
            q[d] = 0;
            u[lay_idx(L, q[0], q[1], q[2])] = w;
        }
        if (p[d] == g1) {
            int q[3] = {p[0], p[1], p[2]};
            q[d] = n[d] + 1;
            u[lay_idx(L, q[0], q[1], q[2])] = w;
        }
    }
}

// Face and edge ghost images of node (ip, jp, kp) with value v, image rules only: along each
// direction the ghosts whose source is this node, then every pair of directions (the composed image
// of the x-then-y-then-z fill when every rule is an image rule).  Corner ghosts are not written: the
// red-black restriction reads the coincident node and the edge points, at most two ghost coordinates.
template <typename T>
__device__ inline void write_images_fe(const Lay &L, T *a, int ip, int jp, int kp, T v, const Bc &bc) {
    const int n[3] = {L.nx, L.ny, L.nz};
    const int p[3] = {ip, jp, kp};
    int g[3][2] = {{0, 0}, {0, 0}, {0, 0}}, cnt[3] = {0, 0, 0};
    T w[3] = {T(1), T(1), T(1)};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int r = bc.rule[d];
        if (r > G_ANTI) continue;
        w[d] = r == G_ANTI ? T(-1) : T(1);
        if (p[d] == ghost_src(r, 0, n[d])) g[d][cnt[d]++] = 0;
        if (p[d] == ghost_src(r, n[d] + 1, n[d])) g[d][cnt[d]++] = n[d] + 1;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
        for (int t = 0; t < cnt[d]; ++t) {
            int q[3] = {p[0], p[1], p[2]};
            q[d] = g[d][t];
            a[lay_idx(L, q[0], q[1], q[2])] = w[d] * v;
        }
#pragma unroll
    for (int d1 = 0; d1 < 2; ++d1)
#pragma unroll
        for (int d2 = d1 + 1; d2 < 3; ++d2)
            for (int t1 = 0; t1 < cnt[d1]; ++t1)
                for (int t2 = 0; t2 < cnt[d2]; ++t2) {
                    int q[3] = {p[0], p[1], p[2]};
                    q[d1] = g[d1][t1];
                    q[d2] = g[d2][t2];
                    a[lay_idx(L, q[0], q[1], q[2])] = w[d1] * w[d2] * v;
                }
}

// ZERO_START: every opposite-colour value is known to be zero (first red half sweep after the level
// was cleared), so the update reduces to -f / sum lc and no u is read.
template <typename T, bool ZERO_START>
__global__ void __launch_bounds__(256) k_rb_sweep(Lay L, Geo G, T *__restrict__ u, const T *__restrict__ f,
                                                  const T *__restrict__ cx, const T *__restrict__ cy,
                                                  const T *__restrict__ cz, int color, Bc bc) {
    unsigned ipu, r;
    if (!geo_slot(G, xcd_block(), ipu, r)) return;
    int ip = (int)ipu;
    unsigned jr = fdivu(r, G.d1, G.inv1);
    int jp = 1 + (int)jr;
    int h = (int)(r - jr * G.d1);
    int kb = (color + 1 + ip + jp) & 1;
    int kp = 2 * h + kb;
    if (kp < 1 || kp > L.nz) return;
    int o = color ^ 1;
    long rc = lay_row(L, ip, jp, color);
    T num;
    const T *X = cx + 4 * ip, *Y = cy + 4 * jp, *Z = cz + 4 * kp;
    T diag = diag3<T>(X, Y, Z);
    if (ZERO_START) {
        num = -f[rc + h];
    } else {
        long ro = lay_row(L, ip, jp, o);
        long sx = 2L * (L.ny + 2) * L.khs;
        long sy = 2L * L.khs;
        T xm = u[ro - sx + h], xp = u[ro + sx + h];
        T ym = u[ro - sy + h], yp = u[ro + sy + h];
        T zm = u[ro + h - 1 + kb], zp = u[ro + h + kb];
        num = X[0] * xp + X[1] * xm + Y[0] * yp + Y[1] * ym + Z[0] * zp + Z[1] * zm - f[rc + h];
    }
    T v = num / diag;
    u[rc + h] = v;
    write_images(L, u, ip, jp, kp, v, bc);
}

// ------------------------------------------------------------------------------------------------
// residual t = f - L u on core nodes (both colours)
// ------------------------------------------------------------------------------------------------
template <typename T, typename TO>
__device__ inline T apply_lap(const Lay &L, const T *__restrict__ u, long row_c, long row_o, int h, int kb,
                              const T *X, const T *Y, const T *Z) {
    long sx = 2L * (L.ny + 2) * L.khs;
    long sy = 2L * L.khs;
    T c = u[row_c + h];
    T xm = u[row_o - sx + h], xp = u[row_o + sx + h];
    T ym = u[row_o - sy + h], yp = u[row_o + sy + h];
    T zm = u[row_o + h - 1 + kb], zp = u[row_o + h + kb];
    return X[0] * xp + X[1] * xm + Y[0] * yp + Y[1] * ym + Z[0] * zp + Z[1] * zm - diag3<T>(X, Y, Z) * c;
}

// color >= 0: that colour's nodes only (geometry geo_sweep), the other colour's entries of t_out
// are left as they are
// IMG: the node's face and edge ghost images by bc are written too (write_images_fe)
template <typename T, typename TO, bool IMG>
__global__ void __launch_bounds__(256) k_residual(Lay L, Geo G, const T *__restrict__ u, const T *__restrict__ f,
                                                  TO *__restrict__ t_out, const T *__restrict__ cx,
                                                  const T *__restrict__ cy, const T *__restrict__ cz, int color,
                                                  Bc bc) {
    unsigned ipu, rr;
    if (!geo_slot(G, xcd_block(), ipu, rr)) return;
    int ip = (int)ipu;
    unsigned jr = fdivu(rr, G.d1, G.inv1);
    unsigned r2 = rr - jr * G.d1;
    int jp = 1 + (int)jr;
    int c = color >= 0 ? color : (r2 >= (unsigned)L.kh ? 1 : 0);
    int h = color >= 0 ? (int)r2 : (int)r2 - c * L.kh;
    int kb = (c + 1 + ip + jp) & 1;
    int kp = 2 * h + kb;
    if (kp < 1 || kp > L.nz) return;
    long rc = lay_row(L, ip, jp, c), ro = lay_row(L, ip, jp, c ^ 1);
    T lu = apply_lap<T, TO>(L, u, rc, ro, h, kb, cx + 4 * ip, cy + 4 * jp, cz + 4 * kp);
    const TO r = (TO)(f[rc + h] - lu);
    t_out[rc + h] = r;
    if (IMG) write_images_fe(L, t_out, ip, jp, kp, r, bc);
}

// ------------------------------------------------------------------------------------------------
// restriction: coarse f(I,J,K) = full weighting of fine t around (2I, 2J, 2K), ghosts included.
// RED0: t vanishes on colour 1 (the residual right after a red-black sweep that ended on colour 1,
// which leaves every colour-1 equation satisfied): the coincident node and the 12 edge points are
// colour 0 (even offset sums), the 6 face and 8 vertex points colour 1, so only the former are read.
// ------------------------------------------------------------------------------------------------
// ZS: the coarse level's first smoothing step, a zero-start colour-0 half sweep (u = -f / diagonal,
// no u read), is done here for the colour-0 nodes from the value just restricted, with its
// face-ghost images (the rest of the level's smoothing then starts with the colour-1 half sweep).
template <typename T, bool RED0, bool ZS>
__global__ void __launch_bounds__(256) k_restrict(Lay F, Lay Cl, Geo G, const T *__restrict__ t, T *__restrict__ fc,
                                                  ZeroStart<T> zs) {
    unsigned ipu, rr;
    if (!geo_slot(G, xcd_block(), ipu, rr)) return;
    int ip = (int)ipu;
    unsigned jr = fdivu(rr, G.d1, G.inv1);
    unsigned r2 = rr - jr * G.d1;
    int jp = 1 + (int)jr;
    int c = r2 >= (unsigned)Cl.kh ? 1 : 0;
    int h = (int)r2 - c * Cl.kh;
    int kp = 2 * h + ((c + 1 + ip + jp) & 1);
    if (kp < 1 || kp > Cl.nz) return;
    // fine padded index of the coincident node: logical 2i -> padded 2(ip-1)+1
    int fi = 2 * ip - 1, fj = 2 * jp - 1, fk = 2 * kp - 1;
#define TT(a, b, e) t[lay_idx(F, fi + (a), fj + (b), fk + (e))]
    if (RED0) {
        T edgePoints = (TT(1, 1, 0) + TT(1, -1, 0) + TT(-1, -1, 0) + TT(-1, 1, 0) + TT(0, 1, 1) + TT(0, -1, 1) +
                        TT(0, -1, -1) + TT(0, 1, -1) + TT(1, 0, 1) + TT(1, 0, -1) + TT(-1, 0, -1) + TT(-1, 0, 1)) *
                       T(0.03125);
        const T val = edgePoints + TT(0, 0, 0) * T(0.125);
        fc[lay_row(Cl, ip, jp, c) + h] = val;
        if (ZS && c == 0) {
            const T v = -val / diag3<T>(zs.cx + 4 * ip, zs.cy + 4 * jp, zs.cz + 4 * kp);
            zs.u[lay_row(Cl, ip, jp, 0) + h] = v;
            write_images(Cl, zs.u, ip, jp, kp, v, zs.bc);
        }
        return;
    }
    T facePoints = (TT(1, 0, 0) + TT(-1, 0, 0) + TT(0, 1, 0) + TT(0, -1, 0) + TT(0, 0, 1) + TT(0, 0, -1)) * T(0.0625);
    T edgePoints = (TT(1, 1, 0) + TT(1, -1, 0) + TT(-1, -1, 0) + TT(-1, 1, 0) + TT(0, 1, 1) + TT(0, -1, 1) +
                    TT(0, -1, -1) + TT(0, 1, -1) + TT(1, 0, 1) + TT(1, 0, -1) + TT(-1, 0, -1) + TT(-1, 0, 1)) *
                   T(0.03125);
    T vertPoints = (TT(1, 1, 1) + TT(1, 1, -1) + TT(1, -1, 1) + TT(-1, 1, 1) + TT(1, -1, -1) + TT(-1, 1, -1) +
                    TT(-1, -1, 1) + TT(-1, -1, -1)) *
                   T(0.015625);
    const T val = facePoints + edgePoints + vertPoints + TT(0, 0, 0) * T(0.125);
    fc[lay_row(Cl, ip, jp, c) + h] = val;
    if (ZS && c == 0) {
        const T v = -val / diag3<T>(zs.cx + 4 * ip, zs.cy + 4 * jp, zs.cz + 4 * kp);
        zs.u[lay_row(Cl, ip, jp, 0) + h] = v;
        write_images(Cl, zs.u, ip, jp, kp, v, zs.bc);
    }
#undef TT
}

// ------------------------------------------------------------------------------------------------
// prolongation: fine u (+)= trilinear interpolation of coarse e (core nodes only).  color >= 0
// updates that colour's nodes only (geometry geo_sweep): a red-black sweep of the other colour
// overwrites those nodes next without reading them, so their correction would be dead work.
// ------------------------------------------------------------------------------------------------
// IMG: the updated node's face-ghost images by bc (image rules only) are written too
template <typename T, typename TF, bool IMG>
__global__ void __launch_bounds__(256) k_prolong(Lay F, Lay Cl, Geo G, const T *__restrict__ e, TF *__restrict__ u,
                                                 int accumulate, int color, Bc bc) {
    unsigned ipu, rr;
    if (!geo_slot(G, xcd_block(), ipu, rr)) return;
    int ip = (int)ipu;
    unsigned jr = fdivu(rr, G.d1, G.inv1);
    unsigned r2 = rr - jr * G.d1;
    int jp = 1 + (int)jr;
    int c = color >= 0 ? color : (r2 >= (unsigned)F.kh ? 1 : 0);
    int h = color >= 0 ? (int)r2 : (int)r2 - c * F.kh;
    int kp = 2 * h + ((c + 1 + ip + jp) & 1);
    if (kp < 1 || kp > F.nz) return;
    int i = ip - 1, j = jp - 1, k = kp - 1;
    int io = i & 1, jo = j & 1, ko = k & 1;
    int ci = i / 2 + 1, cj = j / 2 + 1, ck = k / 2 + 1;   // coarse padded index of i2, j2, k2
#define EE(a, b, q) e[lay_idx(Cl, ci + (a), cj + (b), ck + (q))]
    T v;
    if (io && jo && ko)
        v = (EE(0, 0, 0) + EE(1, 0, 0) + EE(0, 1, 0) + EE(0, 0, 1) + EE(1, 1, 0) + EE(1, 0, 1) + EE(0, 1, 1) +
             EE(1, 1, 1)) / T(8.0);
    else if (io && jo)
        v = (EE(0, 0, 0) + EE(0, 1, 0) + EE(1, 0, 0) + EE(1, 1, 0)) / T(4.0);
    else if (io && ko)
        v = (EE(0, 0, 0) + EE(0, 0, 1) + EE(1, 0, 0) + EE(1, 0, 1)) / T(4.0);
    else if (io)
        v = (EE(0, 0, 0) + EE(1, 0, 0)) / T(2.0);
    else if (jo && ko)
        v = (EE(0, 0, 0) + EE(0, 0, 1) + EE(0, 1, 0) + EE(0, 1, 1)) / T(4.0);
    else if (jo)
        v = (EE(0, 0, 0) + EE(0, 1, 0)) / T(2.0);
    else if (ko)
        v = (EE(0, 0, 0) + EE(0, 0, 1)) / T(2.0);
    else
        v = EE(0, 0, 0);
#undef EE
    long o = lay_row(F, ip, jp, c) + h;
    const TF w = accumulate ? (TF)((TF)u[o] + (TF)v) : (TF)v;
    u[o] = w;
    if (IMG) write_images(F, u, ip, jp, kp, w, bc);
}

// ------------------------------------------------------------------------------------------------
// fills and scalar updates
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_fill(T *__restrict__ a, long n, T v) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) a[t] = v;
}

// a(node) += v over core nodes (core = 1) or over every padded node (core = 0)
template <typename T>
__global__ void __launch_bounds__(256) k_add_scalar(Lay L, Geo G, T *__restrict__ a, const double *__restrict__ dv,
                                                    double scale, int core) {
    unsigned ip, r;
    if (!geo_slot(G, blockIdx.x, ip, r)) return;
    unsigned jp = fdivu(r, G.d1, G.inv1);
    unsigned r2 = r - jp * G.d1;
    int c = r2 >= (unsigned)L.khs ? 1 : 0;
    int h = (int)r2 - c * L.khs;
    if (h >= L.kh) return;
    int kp = 2 * h + ((c + 1 + (int)ip + (int)jp) & 1);
    if (kp > L.nz + 1) return;
    if (core && (ip < 1 || (int)ip > L.nx || jp < 1 || (int)jp > L.ny || kp < 1 || kp > L.nz)) return;
    long t = (long)ip * G.per_plane + r;
    a[t] = (T)((double)a[t] + scale * dv[0]);
}

// ------------------------------------------------------------------------------------------------
// layout conversion between the reference padded array and the split layout
// ------------------------------------------------------------------------------------------------
// core_only = 1: copy core nodes only (ghost slots untouched)
template <typename TS, typename TD>
__global__ void __launch_bounds__(256) k_pack(Lay L, const TS *__restrict__ ref, TD *__restrict__ dst, int core_only) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long NZ = L.nz + 2, NY = L.ny + 2;
    long n = (long)(L.nx + 2) * NY * NZ;
    if (t >= n) return;
    int kp = (int)(t % NZ);
    long r = t / NZ;
    int jp = (int)(r % NY);
    int ip = (int)(r / NY);
    if (core_only && (ip < 1 || ip > L.nx || jp < 1 || jp > L.ny || kp < 1 || kp > L.nz)) return;
    dst[lay_idx(L, ip, jp, kp)] = (TD)ref[t];
}

// Every entry of the split layout in one pass: the core node's value from the padded reference
// array, 0 on ghosts and unused slots (the state of a zero fill followed by a core-only pack).
template <typename TS, typename TD>
__global__ void __launch_bounds__(256) k_pack_full(Lay L, const TS *__restrict__ ref, TD *__restrict__ dst) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lay_size(L)) return;
    long row = t / L.khs;
    int h = (int)(t - row * L.khs);
    int c = (int)(row & 1);
    long col = row >> 1;
    int jp = (int)(col % (L.ny + 2));
    int ip = (int)(col / (L.ny + 2));
    int kp = 2 * h + ((c + 1 + ip + jp) & 1);
    TD v = TD(0);
    if (ip >= 1 && ip <= L.nx && jp >= 1 && jp <= L.ny && kp >= 1 && kp <= L.nz)
        v = (TD)ref[((long)ip * (L.ny + 2) + jp) * (L.nz + 2) + kp];
    dst[t] = v;
}

template <typename TS, typename TD>
__global__ void __launch_bounds__(256) k_unpack(Lay L, const TS *__restrict__ src, TD *__restrict__ ref) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long NZ = L.nz + 2, NY = L.ny + 2;
    long n = (long)(L.nx + 2) * NY * NZ;
    if (t >= n) return;
    int kp = (int)(t % NZ);
    long r = t / NZ;
    int jp = (int)(r % NY);
    int ip = (int)(r / NY);
    ref[t] = (TD)src[lay_idx(L, ip, jp, kp)];
}

// ------------------------------------------------------------------------------------------------
// reductions (deterministic: fixed per-block order, then one block over the partials)
// ------------------------------------------------------------------------------------------------
constexpr int NQ = 4;   // quantities per partial

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// q[0] += sum, q[1] += sum, q[2] = max, q[3] = max over the block; written to part[block*NQ + i]
__device__ inline void block_reduce4(double s0, double s1, double m2, double m3, double *part) {
    __shared__ double sh[4][4];
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    m2 = wave_max(m2);
    m3 = wave_max(m3);
    int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) { sh[w][0] = s0; sh[w][1] = s1; sh[w][2] = m2; sh[w][3] = m3; }
    __syncthreads();
    if (threadIdx.x == 0) {
        int nw = (blockDim.x + 63) >> 6;
        double a = 0, b = 0, c = 0, d = 0;
        for (int i = 0; i < nw; ++i) {
            a += sh[i][0];
            b += sh[i][1];
            c = fmax(c, sh[i][2]);
            d = fmax(d, sh[i][3]);
        }
        double *p = part + (long)blockIdx.x * NQ;
        p[0] = a; p[1] = b; p[2] = c; p[3] = d;
    }
}

// per-block (sum, max) of one value per thread, written at the hardware block index
__device__ inline void block_reduce2(double s, double m, double *part) {
    __shared__ double rs[4], rm[4];
    s = wave_sum(s);
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) {
        rs[threadIdx.x >> 6] = s;
        rm[threadIdx.x >> 6] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0;
        for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) {
            a += rs[w];
            b = fmax(b, rm[w]);
        }
        part[2 * (long)blockIdx.x] = a;
        part[2 * (long)blockIdx.x + 1] = b;
    }
}

// Reduction kernels visit the core nodes of four consecutive x-planes per thread (logical planes
// are plane groups: geo_core4), so one block reduction serves up to 1024 nodes.  color >= 0 visits
// that colour's nodes only (geometry geo_core4c).
template <typename F>
__device__ inline void core4_visit(const Lay &L, const Geo &G, F &&fn, int color = -1) {
    unsigned gu, rr;
    if (!geo_slot(G, xcd_block(), gu, rr)) return;
    unsigned jr = fdivu(rr, G.d1, G.inv1);
    unsigned r2 = rr - jr * G.d1;
    int jp = 1 + (int)jr;
    int c = color >= 0 ? color : (r2 >= (unsigned)L.kh ? 1 : 0);
    int h = color >= 0 ? (int)r2 : (int)r2 - c * L.kh;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        int ip = 1 + 4 * (int)gu + q;
        int kb = (c + 1 + ip + jp) & 1;
        int kp = 2 * h + kb;
        if (ip <= L.nx && kp >= 1 && kp <= L.nz) fn(ip, jp, c, h, kb, kp);
    }
}

// per-cycle residual norm: sum |Lu - f|^2 and max |Lu - f| per block.  The right-hand side terms of computeError do not change during a solve
// and are reduced once per solve (k_f_terms).  color >= 0: over that colour only (after a red-black
// sweep of the other colour, whose residual is then zero).
template <typename T>
__global__ void __launch_bounds__(256) k_norm_res(Lay L, Geo G, const T *__restrict__ u, const T *__restrict__ f,
                                                  const T *__restrict__ cx, const T *__restrict__ cy,
                                                  const T *__restrict__ cz, double *__restrict__ part, int color) {
    double s = 0.0, mm = 0.0;
    core4_visit(L, G, [&](int ip, int jp, int c, int h, int kb, int kp) {
        long rc = lay_row(L, ip, jp, c), ro = lay_row(L, ip, jp, c ^ 1);
        double lu = (double)apply_lap<T, T>(L, u, rc, ro, h, kb, cx + 4 * ip, cy + 4 * jp, cz + 4 * kp);
        double v = fabs(lu - (double)f[rc + h]);
        s += v * v;
        mm = fmax(mm, v);
    }, color);
    block_reduce2(s, mm, part);
}

// the four computeError quantities from the per-block residual partials and the once-per-solve
// right-hand side terms ft = {-, sum f^2 (core), -, max |f| (every padded node)}
__global__ void __launch_bounds__(1024) k_norm_combine(const double *__restrict__ part, int nb,
                                                       const double *__restrict__ ft, double *__restrict__ out) {
    double s = 0.0, m = 0.0;
#pragma unroll 8
    for (int b = (int)threadIdx.x; b < nb; b += 1024) {
        s += part[2 * (long)b];
        m = fmax(m, part[2 * (long)b + 1]);
    }
    __shared__ double rs[16], rm[16];
    s = wave_sum(s);
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) {
        rs[threadIdx.x >> 6] = s;
        rm[threadIdx.x >> 6] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0;
        for (int w = 0; w < 16; ++w) {
            a += rs[w];
            b = fmax(b, rm[w]);
        }
        out[0] = a;
        out[1] = ft[1];
        out[2] = b;
        out[3] = ft[3];
    }
}

// Two-stage final reductions (k_gcr_coef_mb): FINAL_BLOCKS workgroups each reduce a contiguous
// share of the partials in a fixed order into a stage; the workgroup that finishes last (a
// device-scope counter, which it resets) combines the stages in order, so the result does not
// depend on the order in which the workgroups finish.
constexpr int FINAL_BLOCKS = 16;
static_assert(FINAL_BLOCKS * 5 <= FINAL_STAGE, "stage buffer of the two-stage finals");

__device__ inline bool last_block_done(unsigned *counter) {
    __shared__ bool last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(counter, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (last) __threadfence();
    return last;
}


// sum f^2 over the core and max |f| over every padded node
template <typename T>
__global__ void __launch_bounds__(256) k_f_terms(Lay L, Geo G, const T *__restrict__ f, double *__restrict__ part) {
    double s1 = 0, m3 = 0;
    unsigned lb0, lbs, lbe;
    xcd_range(G.total_blocks, lb0, lbs, lbe);
    for (unsigned lb = lb0; lb < lbe; lb += lbs) {
        unsigned ipu, r;
        if (!geo_slot(G, lb, ipu, r)) continue;
        unsigned jpu = fdivu(r, G.d1, G.inv1);
        unsigned r2 = r - jpu * G.d1;
        int c = r2 >= (unsigned)L.khs ? 1 : 0;
        int h = (int)r2 - c * L.khs;
        int ip = (int)ipu, jp = (int)jpu;
        int kp = 2 * h + ((c + 1 + ip + jp) & 1);
        if (h >= L.kh || kp > L.nz + 1) continue;
        double fv = (double)f[(long)ip * G.per_plane + r];
        m3 = fmax(m3, fabs(fv));
        if (ip >= 1 && ip <= L.nx && jp >= 1 && jp <= L.ny && kp >= 1 && kp <= L.nz) s1 += fv * fv;
    }
    block_reduce4(0.0, s1, 0.0, m3, part);
}

// Weighted core sum q0 = sum w_x(i) w_y(j) w_z(k) ref of a reference padded array (the projection's
// left-null-vector weight), read in place.  Work item w covers in-plane entries [256 (w % bpp),
// +256) of the core planes 1 + w / bpp, 1 + w / bpp + groups, ...: each thread's (j, k) and its
// core test are fixed, so the loop over planes is division-free and its loads overlap.  Per-block
// partials {q0, 0, 0, 0}.
__global__ void __launch_bounds__(256) k_wsum_ref(Lay L, unsigned bpp, unsigned groups, float inv2,
                                                  const double *__restrict__ ref, const double *__restrict__ wx,
                                                  const double *__restrict__ wy, const double *__restrict__ wz,
                                                  double *__restrict__ part) {
    const unsigned n1 = (unsigned)L.ny + 2, n2 = (unsigned)L.nz + 2;
    double s0 = 0.0;
    // one pass of the outer loop unless a plane has more than REDUCE_BLOCKS chunks
    for (unsigned w = blockIdx.x; w < bpp * groups; w += gridDim.x) {
        const unsigned chunk = w % bpp, g = w / bpp;
        const unsigned r = chunk * 256u + threadIdx.x;
        const unsigned jj = fdivu(r, n2, inv2), kk = r - jj * n2;
        if (r < n1 * n2 && jj >= 1 && jj <= (unsigned)L.ny && kk >= 1 && kk <= (unsigned)L.nz) {
            const double wjk = wy[jj - 1] * wz[kk - 1];
            const long ps = (long)n1 * n2;
            // the planes 1 + g + q groups, q < nq, in batches of eight loads issued before any is
            // added (the sum runs in plane order, as a plain loop would)
            const unsigned nq = 1 + g <= (unsigned)L.nx ? ((unsigned)L.nx - 1 - g) / groups + 1 : 0;
            unsigned q = 0;
            for (; q + 8 <= nq; q += 8) {
                double v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) v[u] = ref[(1 + g + (q + u) * groups) * ps + r];
#pragma unroll
                for (int u = 0; u < 8; u++) s0 += wx[g + (q + u) * groups] * wjk * v[u];
            }
            for (; q < nq; q++) s0 += wx[g + q * groups] * wjk * ref[(1 + g + q * groups) * ps + r];
        }
    }
    block_reduce4(s0, 0.0, 0.0, 0.0, part);
}

// Every entry of the split layout from a reference padded array in one pass, with the projection's
// shift: core nodes ref + scale q[0], ghosts and unused slots 0 (the state of a full pack followed
// by a core-only shift), and the per-block partials {0, sum v^2 (core), 0, max |v| (core)} of the
// right-hand side terms of computeError.  After the ghost fill of the right-hand side (periodic
// images in x and y, z ghosts kept at 0) the maximum over every padded node equals the core one,
// so these are the terms k_f_terms would reduce.
__global__ void __launch_bounds__(256) k_pack_shift(Lay L, Geo G, const double *__restrict__ ref,
                                                    double *__restrict__ dst, const double *__restrict__ q,
                                                    double scale, double *__restrict__ part) {
    const double sh = scale * q[0];
    double s1 = 0.0, m3 = 0.0;
    unsigned lb0, lbs, lbe;
    xcd_range(G.total_blocks, lb0, lbs, lbe);
    for (unsigned lb = lb0; lb < lbe; lb += lbs) {
        unsigned ipu, r;
        if (!geo_slot(G, lb, ipu, r)) continue;
        const unsigned jpu = fdivu(r, G.d1, G.inv1);
        const unsigned r2 = r - jpu * G.d1;
        const int c = r2 >= (unsigned)L.khs ? 1 : 0;
        const int h = (int)r2 - c * L.khs;
        const int ip = (int)ipu, jp = (int)jpu;
        const int kp = 2 * h + ((c + 1 + ip + jp) & 1);
        double v = 0.0;
        if (h < L.kh && ip >= 1 && ip <= L.nx && jp >= 1 && jp <= L.ny && kp >= 1 && kp <= L.nz) {
            v = ref[((long)ip * (L.ny + 2) + jp) * (L.nz + 2) + kp] + sh;
            s1 += v * v;
            m3 = fmax(m3, fabs(v));
        }
        dst[(long)ip * G.per_plane + r] = v;
    }
    block_reduce4(0.0, s1, 0.0, m3, part);
}

// weighted sums over the reference right-hand side: q0 = sum w_i w_j w_k f (core), q1 = sum f over
// every padded node
template <typename T>
__global__ void __launch_bounds__(256) k_rhs_sums(Lay L, Geo G, const T *__restrict__ f, const double *__restrict__ wx,
                                                  const double *__restrict__ wy, const double *__restrict__ wz,
                                                  double *__restrict__ part) {
    double s0 = 0, s1 = 0;
    unsigned lb0, lbs, lbe;
    xcd_range(G.total_blocks, lb0, lbs, lbe);
    for (unsigned lb = lb0; lb < lbe; lb += lbs) {
        unsigned ipu, r;
        if (!geo_slot(G, lb, ipu, r)) continue;
        unsigned jpu = fdivu(r, G.d1, G.inv1);
        unsigned r2 = r - jpu * G.d1;
        int c = r2 >= (unsigned)L.khs ? 1 : 0;
        int h = (int)r2 - c * L.khs;
        int ip = (int)ipu, jp = (int)jpu;
        int kp = 2 * h + ((c + 1 + ip + jp) & 1);
        if (h >= L.kh || kp > L.nz + 1) continue;
        double fv = (double)f[(long)ip * G.per_plane + r];
        s1 += fv;
        if (ip >= 1 && ip <= L.nx && jp >= 1 && jp <= L.ny && kp >= 1 && kp <= L.nz && wx)
            s0 += wx[ip - 1] * wy[jp - 1] * wz[kp - 1] * fv;
    }
    block_reduce4(s0, s1, 0.0, 0.0, part);
}

__global__ void __launch_bounds__(256) k_reduce_final(const double *__restrict__ part, int nparts, double *__restrict__ out) {
    double s0 = 0, s1 = 0, m2 = 0, m3 = 0;
    // eight partials per thread loaded before any is added (the loads overlap); the sums run in
    // partial order as in a plain strided loop.  m2 and m3 are maxima of absolute values (>= 0).
    const int step = (int)blockDim.x;
    for (int i0 = threadIdx.x; i0 < nparts; i0 += 8 * step) {
        double4 v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int i = i0 + u * step;
            v[u] = i < nparts ? *reinterpret_cast<const double4 *>(part + (long)i * NQ) : make_double4(0, 0, 0, 0);
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            s0 += v[u].x;
            s1 += v[u].y;
            m2 = fmax(m2, v[u].z);
            m3 = fmax(m3, v[u].w);
        }
    }
    // single block: its one "partial" is the result
    block_reduce4(s0, s1, m2, m3, out);
}

// ------------------------------------------------------------------------------------------------
// coarsest level: the whole solve() loop in LDS, one workgroup
// ------------------------------------------------------------------------------------------------
// Every composed ghost value reads core nodes only, so the writes need no ordering among themselves.
template <typename T>
__device__ inline void lds_fill_ghosts(const Lay &L, T *a, const Bc &bc) {
    int NX = L.nx + 2, NY = L.ny + 2, NZ = L.nz + 2;
    int n = NX * NY * NZ;
    const Div3 dv = div3(NY, NZ);
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        int ip, jp, kp;
        split3(dv, t, ip, jp, kp);
        bool ghost = ip == 0 || ip == NX - 1 || jp == 0 || jp == NY - 1 || kp == 0 || kp == NZ - 1;
        T v;
        if (ghost && ghost_resolve(L, a, ip, jp, kp, bc, v)) a[lay_idx(L, ip, jp, kp)] = v;
    }
    __syncthreads();
}

template <typename T>
__device__ inline void lds_sweep(const Lay &L, T *u, const T *f, const T *cx, const T *cy, const T *cz, int color) {
    int n = L.nx * L.ny * L.nz;
    const Div3 dv = div3(L.ny, L.nz);
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        int i, j, k;
        split3(dv, t, i, j, k);
        if (((i + j + k) & 1) != color) continue;
        int ip = i + 1, jp = j + 1, kp = k + 1;
        const T *X = cx + 4 * ip, *Y = cy + 4 * jp, *Z = cz + 4 * kp;
        T num = X[0] * u[lay_idx(L, ip + 1, jp, kp)] + X[1] * u[lay_idx(L, ip - 1, jp, kp)] +
                Y[0] * u[lay_idx(L, ip, jp + 1, kp)] + Y[1] * u[lay_idx(L, ip, jp - 1, kp)] +
                Z[0] * u[lay_idx(L, ip, jp, kp + 1)] + Z[1] * u[lay_idx(L, ip, jp, kp - 1)] - f[lay_idx(L, ip, jp, kp)];
        u[lay_idx(L, ip, jp, kp)] = num / diag3<T>(X, Y, Z);
    }
}

// Loop of poisson_d3.cc:190-247 with the red-black sweep: {ghosts; red; ghosts; black; max|r|}
// until max|r| < tol or more than max_count iterations; ghosts once more at the end.
// The per-iteration count is written to iters[0].
template <typename T>
__global__ void __launch_bounds__(256) k_coarsest(Lay L, T *__restrict__ ug, const T *__restrict__ fg,
                                                  const T *__restrict__ cxg, const T *__restrict__ cyg,
                                                  const T *__restrict__ czg, Bc bc, double tol, int max_count,
                                                  int *__restrict__ iters) {
    extern __shared__ unsigned char smem[];
    long n = lay_size(L);
    T *u = (T *)smem;
    T *f = u + n;
    T *cx = f + n;
    T *cy = cx + 4 * (L.nx + 2);
    T *cz = cy + 4 * (L.ny + 2);
    for (long t = threadIdx.x; t < n; t += blockDim.x) { u[t] = ug[t]; f[t] = fg[t]; }
    for (int t = threadIdx.x; t < 4 * (L.nx + 2); t += blockDim.x) cx[t] = cxg[t];
    for (int t = threadIdx.x; t < 4 * (L.ny + 2); t += blockDim.x) cy[t] = cyg[t];
    for (int t = threadIdx.x; t < 4 * (L.nz + 2); t += blockDim.x) cz[t] = czg[t];
    __syncthreads();
    __shared__ double red[4];
    int it = 0;
    while (true) {
        lds_fill_ghosts(L, u, bc);
        lds_sweep(L, u, f, cx, cy, cz, 0);
        __syncthreads();
        lds_fill_ghosts(L, u, bc);
        lds_sweep(L, u, f, cx, cy, cz, 1);
        __syncthreads();
        double m = -1.0e-10;
        int nc = L.nx * L.ny * L.nz;
        const Div3 dv = div3(L.ny, L.nz);
        for (int t = threadIdx.x; t < nc; t += blockDim.x) {
            int i, j, k;
            split3(dv, t, i, j, k);
            int ip = i + 1, jp = j + 1, kp = k + 1;
            int c = lay_color(ip, jp, kp);
            long rc = lay_row(L, ip, jp, c), ro = lay_row(L, ip, jp, c ^ 1);
            double lu = (double)apply_lap<T, T>(L, u, rc, ro, kp >> 1, kp & 1, cx + 4 * ip, cy + 4 * jp, cz + 4 * kp);
            m = fmax(m, fabs((double)f[rc + (kp >> 1)] - lu));
        }
        m = wave_max(m);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
        __syncthreads();
        double gm = red[0];
        for (int w = 1; w < (int)(blockDim.x + 63) / 64; ++w) gm = fmax(gm, red[w]);
        __syncthreads();
        if (gm < tol) break;
        it += 1;
        if (it > max_count) break;
    }
    lds_fill_ghosts(L, u, bc);
    for (long t = threadIdx.x; t < n; t += blockDim.x) ug[t] = u[t];
    if (threadIdx.x == 0) iters[0] = it;
}

// ------------------------------------------------------------------------------------------------
// Jacobi sweep t = J(u) on core nodes (poisson_d3.cc:152-166)
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_jacobi(Lay L, Geo G, const T *__restrict__ u, const T *__restrict__ f,
                                                T *__restrict__ tout, const T *__restrict__ cx,
                                                const T *__restrict__ cy, const T *__restrict__ cz) {
    unsigned ipu, rr;
    if (!geo_slot(G, xcd_block(), ipu, rr)) return;
    int ip = (int)ipu;
    unsigned jr = fdivu(rr, G.d1, G.inv1);
    unsigned r2 = rr - jr * G.d1;
    int jp = 1 + (int)jr;
    int c = r2 >= (unsigned)L.kh ? 1 : 0;
    int h = (int)r2 - c * L.kh;
    int kb = (c + 1 + ip + jp) & 1;
    int kp = 2 * h + kb;
    if (kp < 1 || kp > L.nz) return;
    long rc = lay_row(L, ip, jp, c), ro = lay_row(L, ip, jp, c ^ 1);
    long sx = 2L * (L.ny + 2) * L.khs, sy = 2L * L.khs;
    const T *X = cx + 4 * ip, *Y = cy + 4 * jp, *Z = cz + 4 * kp;
    T num = X[0] * u[ro + sx + h] + X[1] * u[ro - sx + h] + Y[0] * u[ro + sy + h] + Y[1] * u[ro - sy + h] +
            Z[0] * u[ro + h + kb] + Z[1] * u[ro + h - 1 + kb] - f[rc + h];
    tout[rc + h] = num / diag3<T>(X, Y, Z);
}

// ------------------------------------------------------------------------------------------------
// range projection of a small level's right-hand side (one workgroup): f -= (w.f)/(w.1) on core
// ------------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_project_small(Lay L, T *__restrict__ f, const double *__restrict__ wx,
                                                       const double *__restrict__ wy, const double *__restrict__ wz,
                                                       double wsum) {
    int n = L.nx * L.ny * L.nz;
    double s = 0;
    const Div3 dv = div3(L.ny, L.nz);
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        int i, j, k;
        split3(dv, t, i, j, k);
        s += wx[i] * wy[j] * wz[k] * (double)f[lay_idx(L, i + 1, j + 1, k + 1)];
    }
    __shared__ double red[4];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    double tot = 0;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) tot += red[w];
    double c = tot / wsum;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        int i, j, k;
        split3(dv, t, i, j, k);
        long o = lay_idx(L, i + 1, j + 1, k + 1);
        f[o] = (T)((double)f[o] - c);
    }
}

// ------------------------------------------------------------------------------------------------
// mixed precision: outer FP64 step of iterative refinement
//   CORR: u_out = u_in + e, r = f - L u_out (u_out read on the fly from u_in + e at the stencil)
//   else: u_out untouched, r = f - L u_in
// r goes to the FP32 right-hand side of the inner cycle; the block partials carry
// sum r^2, sum f^2 (core), max |r| (core), max |f| (every padded node) in FP64.
// ------------------------------------------------------------------------------------------------
template <bool CORR>
__global__ void __launch_bounds__(256) k_outer_residual(Lay L, Geo G, const double *__restrict__ u_in,
                                                        const float *__restrict__ e, const double *__restrict__ f,
                                                        double *__restrict__ u_out, float *__restrict__ r32,
                                                        const double *__restrict__ cx, const double *__restrict__ cy,
                                                        const double *__restrict__ cz, Bc bc,
                                                        double *__restrict__ part) {
    double s = 0.0, mm = 0.0;
    const long sx = 2L * (L.ny + 2) * L.khs, sy = 2L * L.khs;
    core4_visit(L, G, [&](int ip, int jp, int c, int h, int kb, int kp) {
        long rc = lay_row(L, ip, jp, c), ro = lay_row(L, ip, jp, c ^ 1);
        auto U = [&](long q) { return CORR ? u_in[q] + (double)e[q] : u_in[q]; };
        double uc = U(rc + h);
        const double *X = cx + 4 * ip, *Y = cy + 4 * jp, *Z = cz + 4 * kp;
        double lu = X[0] * U(ro + sx + h) + X[1] * U(ro - sx + h) + Y[0] * U(ro + sy + h) + Y[1] * U(ro - sy + h) +
                    Z[0] * U(ro + h + kb) + Z[1] * U(ro + h - 1 + kb) - diag3<double>(X, Y, Z) * uc;
        double rv = f[rc + h] - lu;
        r32[rc + h] = (float)rv;
        if (CORR) {
            u_out[rc + h] = uc;
            write_images(L, u_out, ip, jp, kp, uc, bc);
        }
        s += rv * rv;
        mm = fmax(mm, fabs(rv));
    });
    block_reduce2(s, mm, part);
}

// ------------------------------------------------------------------------------------------------
// mixed precision, Krylov-accelerated: flexible GCR(1) preconditioned by one FP32 V-cycle, with the
// FP64 residual recomputed from the iterate every iteration (no recursive residual drift).
//   e      FP32 V-cycle result for the right-hand side r32 (the preconditioned residual)
//   z      previous search direction (FP32; with zero it reduces to a minimal-residual step)
//   q = L e, p = L z;  beta = (q, p) / (p, p);  d = e - beta z;  alpha = (r, L d) / (L d, L d)
//   x_out = x_in + alpha d,  r = f - L x_out
// The directions and the inner products use FP32 data (they only steer the step); the iterate and
// its residual are FP64, so the attained accuracy is that of the FP64 residual.
// ------------------------------------------------------------------------------------------------
// per-block sums (r q, q q, q p, r p, p p), written at the hardware block index.  q and p are
// formed in FP32 with the FP32 operator of the inner cycle (they only steer the step; the iterate's
// residual is recomputed in FP64), the products are summed in FP64.
__device__ inline float lap32(const Lay &L, const float *__restrict__ a, long rc, long ro, int h, int kb,
                              const float *X, const float *Y, const float *Z) {
    const long sx = 2L * (L.ny + 2) * L.khs, sy = 2L * L.khs;
    return X[0] * a[ro + sx + h] + X[1] * a[ro - sx + h] + Y[0] * a[ro + sy + h] + Y[1] * a[ro - sy + h] +
           Z[0] * a[ro + h + kb] + Z[1] * a[ro + h - 1 + kb] - diag3<float>(X, Y, Z) * a[rc + h];
}

// Q1R: e ends on a colour-1 red-black half sweep of level 0 for the right-hand side r32, which
// solves the colour-1 equations: (L e) = r32 there (to FP32 rounding), so q is read, not formed.
template <bool Q1R>
__global__ void __launch_bounds__(256) k_gcr_dots(Lay L, Geo G, const float *__restrict__ e,
                                                  const float *__restrict__ z, const float *__restrict__ r32,
                                                  const float *__restrict__ cx, const float *__restrict__ cy,
                                                  const float *__restrict__ cz, double *__restrict__ part) {
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    core4_visit(L, G, [&](int ip, int jp, int c, int h, int kb, int kp) {
        long rc = lay_row(L, ip, jp, c), ro = lay_row(L, ip, jp, c ^ 1);
        const float *X = cx + 4 * ip, *Y = cy + 4 * jp, *Z = cz + 4 * kp;
        double r = (double)r32[rc + h];
        double q = Q1R && c == 1 ? r : (double)lap32(L, e, rc, ro, h, kb, X, Y, Z);
        double p = (double)lap32(L, z, rc, ro, h, kb, X, Y, Z);
        s[0] += r * q;
        s[1] += q * q;
        s[2] += q * p;
        s[3] += r * p;
        s[4] += p * p;
    });
    __shared__ double sh[4][5];
#pragma unroll
    for (int i = 0; i < 5; i++) s[i] = wave_sum(s[i]);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int i = 0; i < 5; i++) sh[threadIdx.x >> 6][i] = s[i];
    __syncthreads();
    if (threadIdx.x < 5) {
        double a = 0.0;
        for (int w = 0; w < 4; w++) a += sh[w][threadIdx.x];
        part[5 * (long)blockIdx.x + threadIdx.x] = a;
    }
}

// alpha, beta of the step from the block sums; coef = {alpha, beta, (r q), (q q), (q p), (r p), (p p)}.
// When the orthogonalised direction has lost most of its norm to cancellation (or there is no
// previous direction) the step falls back to the minimal-residual step along e.
__device__ inline void gcr_coef_from(const double (&t)[5], double *coef) {
    double rq = t[0], qq = t[1], qp = t[2], rp = t[3], pp = t[4];
    double beta = pp > 0.0 ? qp / pp : 0.0;
    double qn = qq - 2.0 * beta * qp + beta * beta * pp;
    if (!(qn >= 1e-3 * qq)) {
        beta = 0.0;
        qn = qq;
    }
    double alpha = qn > 0.0 ? (rq - beta * rp) / qn : 0.0;
    coef[0] = alpha;
    coef[1] = beta;
    for (int i = 0; i < 5; i++) coef[2 + i] = t[i];
}

__global__ void __launch_bounds__(1024) k_gcr_coef(const double *__restrict__ part, int nb, double *__restrict__ coef) {
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
    for (int b = (int)threadIdx.x; b < nb; b += 1024)
#pragma unroll
        for (int i = 0; i < 5; i++) s[i] += part[5 * (long)b + i];
    __shared__ double sh[16][5];
#pragma unroll
    for (int i = 0; i < 5; i++) s[i] = wave_sum(s[i]);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int i = 0; i < 5; i++) sh[threadIdx.x >> 6][i] = s[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int w = 0; w < 16; w++)
            for (int i = 0; i < 5; i++) t[i] += sh[w][i];
        gcr_coef_from(t, coef);
    }
}

// k_gcr_coef over FINAL_BLOCKS workgroups (two-stage form, see last_block_done)
__global__ void __launch_bounds__(256) k_gcr_coef_mb(const double *__restrict__ part, int nb,
                                                     double *__restrict__ coef, double *__restrict__ stage,
                                                     unsigned *__restrict__ counter) {
    const int per = (nb + (int)gridDim.x - 1) / (int)gridDim.x;
    const int lo = (int)blockIdx.x * per, hi = min(nb, lo + per);
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
    for (int b = lo + (int)threadIdx.x; b < hi; b += 256)
#pragma unroll
        for (int i = 0; i < 5; i++) s[i] += part[5 * (long)b + i];
    __shared__ double sh[4][5];
#pragma unroll
    for (int i = 0; i < 5; i++) s[i] = wave_sum(s[i]);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int i = 0; i < 5; i++) sh[threadIdx.x >> 6][i] = s[i];
    __syncthreads();
    if (threadIdx.x < 5) stage[5 * blockIdx.x + threadIdx.x] = (sh[0][threadIdx.x] + sh[1][threadIdx.x]) +
                                                               (sh[2][threadIdx.x] + sh[3][threadIdx.x]);
    if (!last_block_done(counter)) return;
    if (threadIdx.x == 0) {
        double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int q = 0; q < (int)gridDim.x; q++)
            for (int i = 0; i < 5; i++) t[i] += stage[5 * q + i];
        gcr_coef_from(t, coef);
        atomicExch(counter, 0u);
    }
}

// The step x_out = x_in + alpha (e - beta z) and z_next = e - beta z on core nodes with their
// face-ghost images (3 loads per node).  The FP64 residual of x_out then comes from k_outer_residual
// reading x_out (7 loads per node): measured faster than evaluating x_in + alpha (e - beta z) at the
// seven stencil points inside the step (21 loads per node) although it moves 8 B more per node.
__global__ void __launch_bounds__(256) k_gcr_step(Lay L, Geo G, const double *__restrict__ x_in,
                                                  const float *__restrict__ e, const float *__restrict__ z,
                                                  const double *__restrict__ coef, double *__restrict__ x_out,
                                                  float *__restrict__ z_next, Bc bc, Bc ebc) {
    const double alpha = coef[0], beta = coef[1];
    core4_visit(L, G, [&](int ip, int jp, int c, int h, int kb, int kp) {
        (void)kb;
        const long q = lay_row(L, ip, jp, c) + h;
        double dc = (double)e[q] - beta * (double)z[q];
        double uc = x_in[q] + alpha * dc;
        x_out[q] = uc;
        write_images(L, x_out, ip, jp, kp, uc, bc);
        float dz = (float)dc;
        z_next[q] = dz;
        write_images(L, z_next, ip, jp, kp, dz, ebc);
    });
}

// ------------------------------------------------------------------------------------------------
// the coarse end of the V-cycle in one workgroup: levels first..last of the down leg, the coarsest
// solve and the up leg back to `first`, with every level's u and f held in LDS
// (poisson.cc:239-279 restricted to those levels; poisson_d3.cc smooth/coarsen/solve/prolong)
// ------------------------------------------------------------------------------------------------
template <typename T>
__device__ inline void blk_sweep(const Lay &L, T *u, const T *f, const T *cx, const T *cy, const T *cz, int color,
                                 const Bc &bc, bool zero_start) {
    int n = L.nx * L.ny * L.kh;
    const Div3 dv = div3(L.ny, L.kh);
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        int ip, jp, h;
        split3(dv, t, ip, jp, h);
        ip += 1;
        jp += 1;
        int kb = (color + 1 + ip + jp) & 1;
        int kp = 2 * h + kb;
        if (kp < 1 || kp > L.nz) continue;
        long rc = lay_row(L, ip, jp, color);
        const T *X = cx + 4 * ip, *Y = cy + 4 * jp, *Z = cz + 4 * kp;
        T num;
        if (zero_start) {
            num = -f[rc + h];
        } else {
            long ro = lay_row(L, ip, jp, color ^ 1);
            long sx = 2L * (L.ny + 2) * L.khs, sy = 2L * L.khs;
            num = X[0] * u[ro + sx + h] + X[1] * u[ro - sx + h] + Y[0] * u[ro + sy + h] + Y[1] * u[ro - sy + h] +
                  Z[0] * u[ro + h + kb] + Z[1] * u[ro + h - 1 + kb] - f[rc + h];
        }
        T v = num / diag3<T>(X, Y, Z);
        u[rc + h] = v;
        write_images(L, u, ip, jp, kp, v, bc);
    }
    __syncthreads();
}

template <typename T>
__device__ inline void blk_zero(const Lay &L, T *u) {
    long n = lay_size(L);
    for (long t = threadIdx.x; t < n; t += blockDim.x) u[t] = T(0);
    __syncthreads();
}

// Residual of level L at padded node (ip, jp, kp) as the restriction reads it (residual-pad rules
// tb).  The position is resolved without branches: a ghost position is mapped to the core node
// whose residual it carries (periodic / mirror / antisymmetric image, factor +-1) or given factor 0
// (zero pads), and the core residual f - L u (the expression of apply_lap) is evaluated there, so
// the 27 evaluations of a restriction can overlap their loads.
template <typename T>
__device__ inline T blk_res_src(const Lay &L, const T *u, const T *f, const T *cx, const T *cy, const T *cz,
                                const Bc &tb, int ip, int jp, int kp) {
    const int n[3] = {L.nx, L.ny, L.nz};
    int q[3] = {ip, jp, kp};
    int fac = 1;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int g = q[d], r = tb.rule[d];
        const bool ghost = g == 0 || g == n[d] + 1;
        const bool image = r <= G_ANTI;
        fac = !ghost ? fac : !image ? 0 : r == G_ANTI ? -fac : fac;
        q[d] = !ghost ? g : image ? ghost_src(r, g, n[d]) : 1;
    }
    const int sy = 2 * L.khs, sx = (L.ny + 2) * sy;
    const int c = lay_color(q[0], q[1], q[2]);
    const int base = (q[0] * (L.ny + 2) + q[1]) * sy;
    const int rc = base + c * L.khs, ro = base + (c ^ 1) * L.khs;
    const int h = q[2] >> 1, kb = q[2] & 1;
    const T *X = cx + 4 * q[0], *Y = cy + 4 * q[1], *Z = cz + 4 * q[2];
    T lu = X[0] * u[ro + sx + h] + X[1] * u[ro - sx + h] + Y[0] * u[ro + sy + h] + Y[1] * u[ro - sy + h] +
           Z[0] * u[ro + h + kb] + Z[1] * u[ro + h - 1 + kb] - diag3<T>(X, Y, Z) * u[rc + h];
    T res = f[rc + h] - lu;
    return fac == 0 ? T(0) : fac < 0 ? -res : res;
}

// 27-point full weighting (poisson_d3.cc:250-298) of the residuals R(a, b, e) at the fine offsets
// a, b, e in {-1, 0, 1} around a coarse node's fine position: faces 1/16, edges 1/32, vertices
// 1/64, centre 1/8, summed in this order
template <typename T, typename F>
__device__ inline T full_weight27(F &&R) {
    T facePoints = (R(1, 0, 0) + R(-1, 0, 0) + R(0, 1, 0) + R(0, -1, 0) + R(0, 0, 1) + R(0, 0, -1)) * T(0.0625);
    T edgePoints = (R(1, 1, 0) + R(1, -1, 0) + R(-1, -1, 0) + R(-1, 1, 0) + R(0, 1, 1) + R(0, -1, 1) + R(0, -1, -1) +
                    R(0, 1, -1) + R(1, 0, 1) + R(1, 0, -1) + R(-1, 0, -1) + R(-1, 0, 1)) *
                   T(0.03125);
    T vertPoints = (R(1, 1, 1) + R(1, 1, -1) + R(1, -1, 1) + R(-1, 1, 1) + R(1, -1, -1) + R(-1, 1, -1) + R(-1, -1, 1) +
                    R(-1, -1, -1)) *
                   T(0.015625);
    return facePoints + edgePoints + vertPoints + R(0, 0, 0) * T(0.125);
}

// one thread per coarse node, its 27 fine residuals evaluated in turn
template <typename T>
__device__ inline void blk_restrict(const Lay &F, const T *uf, const T *ff, const T *cx, const T *cy, const T *cz,
                                    const Bc &tb, const Lay &Cl, T *fc) {
    int n = Cl.nx * Cl.ny * Cl.nz;
    const Div3 dv = div3(Cl.ny, Cl.nz);
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        int i, j, k;
        split3(dv, t, i, j, k);
        int fi = 2 * i + 1, fj = 2 * j + 1, fk = 2 * k + 1;
        fc[lay_idx(Cl, i + 1, j + 1, k + 1)] = full_weight27<T>([&](int a, int b, int e) {
            return blk_res_src<T>(F, uf, ff, cx, cy, cz, tb, fi + a, fj + b, fk + e);
        });
    }
    __syncthreads();
}

// The same restriction with one thread per (coarse node, tap): rounds of kTapNodes coarse nodes,
// whose 27 x kTapNodes fine residuals are evaluated at once into scr, then weighted by one thread
// per node with the same arithmetic as blk_restrict.  For the smallest levels, where one thread
// per coarse node leaves most of the workgroup idle behind 27 dependent residual evaluations.
constexpr int kTapNodes = 37;
constexpr int kTapScratch = 27 * kTapNodes;

template <typename T>
__device__ inline void blk_restrict_taps(const Lay &F, const T *uf, const T *ff, const T *cx, const T *cy,
                                         const T *cz, const Bc &tb, const Lay &Cl, T *fc, T *scr) {
    const int n = Cl.nx * Cl.ny * Cl.nz;
    const Div3 dv = div3(Cl.ny, Cl.nz);
    const int t = threadIdx.x;
    for (int first = 0; first < n; first += kTapNodes) {
        if (t < kTapScratch) {
            int node = first + t / 27, tap = t % 27;
            if (node < n) {
                int i, j, k;
                split3(dv, node, i, j, k);
                scr[t] = blk_res_src<T>(F, uf, ff, cx, cy, cz, tb, 2 * i + tap / 9, 2 * j + (tap / 3) % 3,
                                       2 * k + tap % 3);
            }
        }
        __syncthreads();
        if (t < kTapNodes && first + t < n) {
            const T *R = scr + 27 * t;
            int i, j, k;
            split3(dv, first + t, i, j, k);
            fc[lay_idx(Cl, i + 1, j + 1, k + 1)] =
                full_weight27<T>([&](int a, int b, int e) { return R[(a + 1) * 9 + (b + 1) * 3 + (e + 1)]; });
        }
        __syncthreads();
    }
}

// The red-black forms of the two restrictions above: after a red-black sweep that ended on colour 1
// the colour-1 residual is zero, and of the 27 points only the coincident node and the 12 edge
// points are colour 0 (k_restrict<T, true>): 13 residual evaluations per coarse node, weighted in
// k_restrict<T, true>'s order.  kRbTaps lists the 12 edge offsets in that order, then the centre.
__constant__ signed char kRbTaps[13][3] = {{1, 1, 0},  {1, -1, 0}, {-1, -1, 0}, {-1, 1, 0}, {0, 1, 1},
                                           {0, -1, 1}, {0, -1, -1}, {0, 1, -1}, {1, 0, 1},  {1, 0, -1},
                                           {-1, 0, -1}, {-1, 0, 1}, {0, 0, 0}};
constexpr int kRbTapNodes = kTapScratch / 13;

template <typename T, typename R>
__device__ inline T rb_weight13(R &&res) {
    T edgePoints = (res(0) + res(1) + res(2) + res(3) + res(4) + res(5) + res(6) + res(7) + res(8) + res(9) +
                    res(10) + res(11)) *
                   T(0.03125);
    return edgePoints + res(12) * T(0.125);
}

template <typename T>
__device__ inline void blk_restrict_rb(const Lay &F, const T *uf, const T *ff, const T *cx, const T *cy, const T *cz,
                                       const Bc &tb, const Lay &Cl, T *fc) {
    int n = Cl.nx * Cl.ny * Cl.nz;
    const Div3 dv = div3(Cl.ny, Cl.nz);
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        int i, j, k;
        split3(dv, t, i, j, k);
        int fi = 2 * i + 1, fj = 2 * j + 1, fk = 2 * k + 1;
        fc[lay_idx(Cl, i + 1, j + 1, k + 1)] = rb_weight13<T>([&](int q) {
            return blk_res_src<T>(F, uf, ff, cx, cy, cz, tb, fi + kRbTaps[q][0], fj + kRbTaps[q][1],
                                  fk + kRbTaps[q][2]);
        });
    }
    __syncthreads();
}

template <typename T>
__device__ inline void blk_restrict_rb_taps(const Lay &F, const T *uf, const T *ff, const T *cx, const T *cy,
                                            const T *cz, const Bc &tb, const Lay &Cl, T *fc, T *scr) {
    const int n = Cl.nx * Cl.ny * Cl.nz;
    const Div3 dv = div3(Cl.ny, Cl.nz);
    const int t = threadIdx.x;
    for (int first = 0; first < n; first += kRbTapNodes) {
        if (t < 13 * kRbTapNodes) {
            int node = first + t / 13, tap = t % 13;
            if (node < n) {
                int i, j, k;
                split3(dv, node, i, j, k);
                scr[t] = blk_res_src<T>(F, uf, ff, cx, cy, cz, tb, 2 * i + 1 + kRbTaps[tap][0],
                                        2 * j + 1 + kRbTaps[tap][1], 2 * k + 1 + kRbTaps[tap][2]);
            }
        }
        __syncthreads();
        if (t < kRbTapNodes && first + t < n) {
            const T *R = scr + 13 * t;
            int i, j, k;
            split3(dv, first + t, i, j, k);
            fc[lay_idx(Cl, i + 1, j + 1, k + 1)] = rb_weight13<T>([&](int q) { return R[q]; });
        }
        __syncthreads();
    }
}

// uf += P e on the core nodes of the fine level (trilinear, poisson_d3.cc:301-388).  The items
// t, t + blockDim, t + 2 blockDim of a thread update distinct nodes, so their loads are all issued
// before the first update is stored.  With `images` every updated node also rewrites its face-ghost
// images (all smoothing ghost rules are images), which replaces a ghost fill of the whole level.
constexpr int kBlkBatch = 3;

template <typename T>
__device__ inline void blk_prolong_add(const Lay &F, T *uf, const Lay &Cl, const T *e, const Bc &bc, bool images,
                                       int color = -1) {
    // color 1: the colour-1 nodes only (slots (i, j, h), kp = 2h + parity), the post-smoothing's first
    // colour-0 half sweep overwrites the colour-0 nodes without reading them
    const int kc = color == 1 ? (F.nz + 2) / 2 : F.nz;
    const int n = F.nx * F.ny * kc;
    const Div3 dv = div3(F.ny, kc);
    for (int t0 = threadIdx.x; t0 < n; t0 += kBlkBatch * blockDim.x) {
        T val[kBlkBatch];
        int pi[kBlkBatch], pj[kBlkBatch], pk[kBlkBatch];
        bool ok[kBlkBatch];
#pragma unroll
        for (int b = 0; b < kBlkBatch; b++) {
            const int t = t0 + b * (int)blockDim.x;
            ok[b] = false;
            if (t >= n) continue;
            int i, j, k;
            split3(dv, t, i, j, k);
            if (color == 1) {
                const int kp = 2 * k + ((i + j) & 1);
                if (kp < 1 || kp > F.nz) continue;
                k = kp - 1;
            }
            ok[b] = true;
            int io = i & 1, jo = j & 1, ko = k & 1;
            int ci = i / 2 + 1, cj = j / 2 + 1, ck = k / 2 + 1;
#define EE(a, b2, q) e[lay_idx(Cl, ci + (a), cj + (b2), ck + (q))]
            T v;
            if (io && jo && ko)
                v = (EE(0, 0, 0) + EE(1, 0, 0) + EE(0, 1, 0) + EE(0, 0, 1) + EE(1, 1, 0) + EE(1, 0, 1) + EE(0, 1, 1) +
                     EE(1, 1, 1)) / T(8.0);
            else if (io && jo)
                v = (EE(0, 0, 0) + EE(0, 1, 0) + EE(1, 0, 0) + EE(1, 1, 0)) / T(4.0);
            else if (io && ko)
                v = (EE(0, 0, 0) + EE(0, 0, 1) + EE(1, 0, 0) + EE(1, 0, 1)) / T(4.0);
            else if (io)
                v = (EE(0, 0, 0) + EE(1, 0, 0)) / T(2.0);
            else if (jo && ko)
                v = (EE(0, 0, 0) + EE(0, 0, 1) + EE(0, 1, 0) + EE(0, 1, 1)) / T(4.0);
            else if (jo)
                v = (EE(0, 0, 0) + EE(0, 1, 0)) / T(2.0);
            else if (ko)
                v = (EE(0, 0, 0) + EE(0, 0, 1)) / T(2.0);
            else
                v = EE(0, 0, 0);
#undef EE
            pi[b] = i + 1;
            pj[b] = j + 1;
            pk[b] = k + 1;
            val[b] = uf[lay_idx(F, i + 1, j + 1, k + 1)] + v;
        }
#pragma unroll
        for (int b = 0; b < kBlkBatch; b++) {
            if (!ok[b]) continue;
            uf[lay_idx(F, pi[b], pj[b], pk[b])] = val[b];
            if (images) write_images(F, uf, pi[b], pj[b], pk[b], val[b], bc);
        }
    }
    __syncthreads();
}

template <typename T>
__device__ inline double blk_max_res(const Lay &L, const T *u, const T *f, const T *cx, const T *cy, const T *cz,
                                     double *red) {
    double m = -1.0e-10;
    int nc = L.nx * L.ny * L.nz;
    const Div3 dv = div3(L.ny, L.nz);
    for (int t = threadIdx.x; t < nc; t += blockDim.x) {
        int i, j, k;
        split3(dv, t, i, j, k);
        int ip = i + 1, jp = j + 1, kp = k + 1;
        int c = lay_color(ip, jp, kp);
        long rc = lay_row(L, ip, jp, c), ro = lay_row(L, ip, jp, c ^ 1);
        double lu = (double)apply_lap<T, T>(L, u, rc, ro, kp >> 1, kp & 1, cx + 4 * ip, cy + 4 * jp, cz + 4 * kp);
        m = fmax(m, fabs((double)f[rc + (kp >> 1)] - lu));
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    double gm = red[0];
    for (int w = 1; w < (int)(blockDim.x + 63) / 64; ++w) gm = fmax(gm, red[w]);
    __syncthreads();
    return gm;
}

template <typename T>
__device__ inline void blk_project(const Lay &L, T *f, const double *wx, const double *wy, const double *wz,
                                   double wsum, double *red) {
    int n = L.nx * L.ny * L.nz;
    double s = 0;
    const Div3 dv = div3(L.ny, L.nz);
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        int i, j, k;
        split3(dv, t, i, j, k);
        s += wx[i] * wy[j] * wz[k] * (double)f[lay_idx(L, i + 1, j + 1, k + 1)];
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    double tot = 0;
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) tot += red[w];
    __syncthreads();
    double c = tot / wsum;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        int i, j, k;
        split3(dv, t, i, j, k);
        long o = lay_idx(L, i + 1, j + 1, k + 1);
        f[o] = (T)((double)f[o] - c);
    }
    __syncthreads();
}

// The coarsest-level iteration of the coarse cycle in one wavefront: the coarsest level has a few
// dozen nodes, so one wave updates them with wave-level synchronisation instead of workgroup
// barriers.  Same red-black order, same ghost composition, same stop rule as the whole-block path.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

template <typename T>
__device__ inline void wv_fill_ghosts(const Lay &L, T *a, const Bc &bc, int lane) {
    int NX = L.nx + 2, NY = L.ny + 2, NZ = L.nz + 2;
    int n = NX * NY * NZ;
    const Div3 dv = div3(NY, NZ);
    for (int t = lane; t < n; t += 64) {
        int ip, jp, kp;
        split3(dv, t, ip, jp, kp);
        bool ghost = ip == 0 || ip == NX - 1 || jp == 0 || jp == NY - 1 || kp == 0 || kp == NZ - 1;
        T v;
        if (ghost && ghost_resolve(L, a, ip, jp, kp, bc, v)) a[lay_idx(L, ip, jp, kp)] = v;
    }
    wave_sync();
}

template <typename T>
__device__ inline void wv_sweep(const Lay &L, T *u, const T *f, const T *cx, const T *cy, const T *cz, int color,
                                int lane) {
    int n = L.nx * L.ny * L.nz;
    const Div3 dv = div3(L.ny, L.nz);
    for (int t = lane; t < n; t += 64) {
        int i, j, k;
        split3(dv, t, i, j, k);
        if (((i + j + k) & 1) != color) continue;
        int ip = i + 1, jp = j + 1, kp = k + 1;
        const T *X = cx + 4 * ip, *Y = cy + 4 * jp, *Z = cz + 4 * kp;
        T num = X[0] * u[lay_idx(L, ip + 1, jp, kp)] + X[1] * u[lay_idx(L, ip - 1, jp, kp)] +
                Y[0] * u[lay_idx(L, ip, jp + 1, kp)] + Y[1] * u[lay_idx(L, ip, jp - 1, kp)] +
                Z[0] * u[lay_idx(L, ip, jp, kp + 1)] + Z[1] * u[lay_idx(L, ip, jp, kp - 1)] - f[lay_idx(L, ip, jp, kp)];
        u[lay_idx(L, ip, jp, kp)] = num / diag3<T>(X, Y, Z);
    }
    wave_sync();
}

template <typename T>
__device__ inline double wv_max_res(const Lay &L, const T *u, const T *f, const T *cx, const T *cy, const T *cz,
                                    int lane) {
    double m = -1.0e-10;
    int nc = L.nx * L.ny * L.nz;
    const Div3 dv = div3(L.ny, L.nz);
    for (int t = lane; t < nc; t += 64) {
        int i, j, k;
        split3(dv, t, i, j, k);
        int ip = i + 1, jp = j + 1, kp = k + 1;
        int c = lay_color(ip, jp, kp);
        long rc = lay_row(L, ip, jp, c), ro = lay_row(L, ip, jp, c ^ 1);
        double lu = (double)apply_lap<T, T>(L, u, rc, ro, kp >> 1, kp & 1, cx + 4 * ip, cy + 4 * jp, cz + 4 * kp);
        m = fmax(m, fabs((double)f[rc + (kp >> 1)] - lu));
    }
    return wave_max(m);
}

// The same coarsest-level iteration for a level of at most 64 core nodes and 256 padded entries:
// each lane keeps one node's neighbour indices, coefficients and right-hand side in registers, and
// up to four ghost entries with where they take their value, so an iteration is a few LDS reads
// per lane instead of the index arithmetic and ghost composition of every pass.  The expressions
// are those of wv_sweep and wv_max_res, and the order of the reference loop is kept: the ghosts
// whose source has colour c are refilled after the colour-c half sweep, which is where the next
// imposeBC would first change them, and the colour-1 ones only after the residual, which the
// reference evaluates before its next imposeBC (poisson_d3.cc:190-247).  Returns the iteration count.
template <typename T>
__device__ inline int wv_coarsest_small(const Lay &L, T *u, const T *f, const T *cx, const T *cy, const T *cz,
                                        const Bc &bc, double tol, int max_count, int lane) {
    const int nc = L.nx * L.ny * L.nz;
    const bool own = lane < nc;
    long me = 0, nb[6] = {0, 0, 0, 0, 0, 0};
    T co[6] = {T(0), T(0), T(0), T(0), T(0), T(0)}, dg = T(1), fm = T(0);
    int col = 2;
    if (own) {
        int k = lane % L.nz, r = lane / L.nz;
        int j = r % L.ny, i = r / L.ny;
        int ip = i + 1, jp = j + 1, kp = k + 1;
        me = lay_idx(L, ip, jp, kp);
        nb[0] = lay_idx(L, ip + 1, jp, kp);
        nb[1] = lay_idx(L, ip - 1, jp, kp);
        nb[2] = lay_idx(L, ip, jp + 1, kp);
        nb[3] = lay_idx(L, ip, jp - 1, kp);
        nb[4] = lay_idx(L, ip, jp, kp + 1);
        nb[5] = lay_idx(L, ip, jp, kp - 1);
        const T *X = cx + 4 * ip, *Y = cy + 4 * jp, *Z = cz + 4 * kp;
        co[0] = X[0]; co[1] = X[1]; co[2] = Y[0]; co[3] = Y[1]; co[4] = Z[0]; co[5] = Z[1];
        dg = diag3<T>(X, Y, Z);
        fm = f[me];
        col = (i + j + k) & 1;
    }
    const int NX = L.nx + 2, NY = L.ny + 2, NZ = L.nz + 2;
    const int np = NX * NY * NZ;
    long gd[4] = {0, 0, 0, 0}, gs[4] = {0, 0, 0, 0};
    T gsg[4] = {T(1), T(1), T(1), T(1)}, gc[4] = {T(0), T(0), T(0), T(0)};
    int gk[4] = {0, 0, 0, 0}, gcol[4] = {0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        int p = lane + 64 * q;
        if (p >= np) continue;
        int kp = p % NZ, r = p / NZ;
        int jp = r % NY, ip = r / NY;
        if (!(ip == 0 || ip == NX - 1 || jp == 0 || jp == NY - 1 || kp == 0 || kp == NZ - 1)) continue;
        gk[q] = ghost_source<T>(L, ip, jp, kp, bc, gs[q], gsg[q], gc[q]);
        gd[q] = lay_idx(L, ip, jp, kp);
        if (gk[q] == 1) {
            // colour of the source node (lay_idx's row carries it)
            long row = gs[q] / L.khs;
            gcol[q] = (int)(row & 1);
        }
    }
    // colour c < 2: the ghosts whose source has colour c; c == 2: every ghost
    auto fill = [&](int c) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (gk[q] == 1 && (c == 2 || gcol[q] == c)) u[gd[q]] = gsg[q] * u[gs[q]];
            else if (gk[q] == 2) u[gd[q]] = gc[q];
        }
    };
    fill(2);
    wave_sync();
    int it = 0;
    while (true) {
        if (col == 0)
            u[me] = (co[0] * u[nb[0]] + co[1] * u[nb[1]] + co[2] * u[nb[2]] + co[3] * u[nb[3]] + co[4] * u[nb[4]] +
                     co[5] * u[nb[5]] - fm) / dg;
        wave_sync();
        fill(0);
        wave_sync();
        if (col == 1)
            u[me] = (co[0] * u[nb[0]] + co[1] * u[nb[1]] + co[2] * u[nb[2]] + co[3] * u[nb[3]] + co[4] * u[nb[4]] +
                     co[5] * u[nb[5]] - fm) / dg;
        wave_sync();
        double m = -1.0e-10;
        if (own) {
            T c = u[me];
            T lu = co[0] * u[nb[0]] + co[1] * u[nb[1]] + co[2] * u[nb[2]] + co[3] * u[nb[3]] + co[4] * u[nb[4]] +
                   co[5] * u[nb[5]] - dg * c;
            m = fmax(m, fabs((double)fm - (double)lu));
        }
        const double gm = wave_max(m);
        fill(1);
        wave_sync();
        if (gm < tol) break;
        it += 1;
        if (it > max_count) break;
    }
    return it;
}

// u = P f on the core nodes of the coarsest level (P transposed: pt[j * nc + i] = P[i][j]), node
// order (i, j, k) with k fastest; ghosts are left to the caller.  FP64 accumulation.
template <typename T>
__device__ inline void blk_direct(const Lay &L, T *u, const T *f, const T *__restrict__ pt, int nc) {
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        double acc = 0.0;
        // eight entries of the operator's column loaded before any is used (they come from global
        // memory); the products are added in node order
        for (int j0 = 0; j0 < nc; j0 += 8) {
            double p[8];
            T fv[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int j = j0 + q;
                p[q] = j < nc ? (double)pt[(long)j * nc + i] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int j = j0 + q < nc ? j0 + q : nc - 1;
                const int c = j % L.nz, r = j / L.nz;
                fv[q] = f[lay_idx(L, 1 + r / L.ny, 1 + r % L.ny, 1 + c)];
            }
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (j0 + q < nc) acc += p[q] * (double)fv[q];
        }
        int k = i % L.nz, r = i / L.nz;
        u[lay_idx(L, 1 + r / L.ny, 1 + r % L.ny, 1 + k)] = (T)acc;
    }
    __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(1024) k_coarse_cycle(CoarseDesc D, T *__restrict__ u_first, const T *__restrict__ f_first,
                                                       int *__restrict__ iters) {
    extern __shared__ unsigned char smem[];
    T *base = (T *)smem;
    __shared__ double red[16];
    const int nl = D.last - D.first + 1;
    // the coarsest level's dense operator (exact coarsest solve), one entry per thread, loaded now and
    // placed in the restriction scratch for the coarsest phase, where it fits
    const int nc2 = D.pinv ? D.ncd * D.ncd : 0;
    const bool pt_lds = D.off_s >= 0 && nc2 > 0 && nc2 <= kTapScratch;
    const T ptv = pt_lds && (int)threadIdx.x < nc2 ? ((const T *)D.pinv)[threadIdx.x] : T(0);
    // coefficients and the right-hand side of the first level into LDS; u(first) starts at zero.
    // The global loads are issued in batches before their LDS stores (each is a round trip to L2
    // or memory; the copies of one workgroup would otherwise run one after another).
    {
        T cv[8][3];
#pragma unroll
        for (int q = 0; q < 8; q++)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const int cnt = q < nl ? 4 * ((a == 0 ? D.L[q].nx : a == 1 ? D.L[q].ny : D.L[q].nz) + 2) : 0;
                cv[q][a] = (int)threadIdx.x < cnt ? ((const T *)D.cg[q][a])[threadIdx.x] : T(0);
            }
#pragma unroll
        for (int q = 0; q < 8; q++)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const int cnt = q < nl ? 4 * ((a == 0 ? D.L[q].nx : a == 1 ? D.L[q].ny : D.L[q].nz) + 2) : 0;
                if ((int)threadIdx.x < cnt) base[D.off_c[q][a] + threadIdx.x] = cv[q][a];
                // arrays longer than the workgroup (not met by the coarse levels) copy the rest
                for (int t = (int)threadIdx.x + (int)blockDim.x; t < cnt; t += blockDim.x)
                    base[D.off_c[q][a] + t] = ((const T *)D.cg[q][a])[t];
            }
    }
    {
        const long n = lay_size(D.L[0]);
        T *f = base + D.off_f[0];
        const long step = blockDim.x;
        for (long t0 = threadIdx.x; t0 < n; t0 += 8 * step) {
            T v[8];
#pragma unroll
            for (int q = 0; q < 8; q++) v[q] = t0 + q * step < n ? f_first[t0 + q * step] : T(0);
#pragma unroll
            for (int q = 0; q < 8; q++)
                if (t0 + q * step < n) f[t0 + q * step] = v[q];
        }
    }
    __syncthreads();
#define CX(q) (base + D.off_c[q][0])
#define CY(q) (base + D.off_c[q][1])
#define CZ(q) (base + D.off_c[q][2])
#define UU(q) (base + D.off_u[q])
#define FF(q) (base + D.off_f[q])
    // down leg
    for (int q = 0; q + 1 < nl; q++) {
        const Lay &L = D.L[q];
        for (int s = 0; s < D.pre; s++) {
            blk_sweep<T>(L, UU(q), FF(q), CX(q), CY(q), CZ(q), 0, D.bc, s == 0);
            blk_sweep<T>(L, UU(q), FF(q), CX(q), CY(q), CZ(q), 1, D.bc, false);
        }
        if (D.pre == 0) {
            blk_zero<T>(L, UU(q));
        }
        const Lay &C = D.L[q + 1];
        if (D.pre > 0) {
            // the pre-smoothing ended on colour 1: only the colour-0 residual points are restricted
            if (D.off_s >= 0 && C.nx * C.ny * C.nz <= 4 * kRbTapNodes)
                blk_restrict_rb_taps<T>(L, UU(q), FF(q), CX(q), CY(q), CZ(q), D.tbc, C, FF(q + 1), base + D.off_s);
            else
                blk_restrict_rb<T>(L, UU(q), FF(q), CX(q), CY(q), CZ(q), D.tbc, C, FF(q + 1));
        } else if (D.off_s >= 0 && C.nx * C.ny * C.nz <= D.tap_max) {
            blk_restrict_taps<T>(L, UU(q), FF(q), CX(q), CY(q), CZ(q), D.tbc, C, FF(q + 1), base + D.off_s);
        } else {
            blk_restrict<T>(L, UU(q), FF(q), CX(q), CY(q), CZ(q), D.tbc, C, FF(q + 1));
        }
    }
    // coarsest solve (poisson_d3.cc:190-247 with the red-black sweep)
    {
        int q = nl - 1;
        const Lay &L = D.L[q];
        blk_zero<T>(L, UU(q));
        if (D.pinv) {
            const T *pt = (const T *)D.pinv;
            if (pt_lds) {
                if ((int)threadIdx.x < nc2) base[D.off_s + threadIdx.x] = ptv;
                __syncthreads();
                pt = base + D.off_s;
            }
            blk_direct<T>(L, UU(q), FF(q), pt, D.ncd);
            lds_fill_ghosts(L, UU(q), D.bc);
            if (threadIdx.x == 0) iters[0] = 0;
        } else {
        if (D.project) blk_project<T>(L, FF(q), D.wc[0], D.wc[1], D.wc[2], D.wcsum, red);
        const bool small = L.nx * L.ny * L.nz <= 64 && (L.nx + 2) * (L.ny + 2) * (L.nz + 2) <= 256;
        if (small && threadIdx.x < 64) {
            const int it = wv_coarsest_small<T>(L, UU(q), FF(q), CX(q), CY(q), CZ(q), D.bc, D.tol, D.max_count,
                                                (int)threadIdx.x);
            if (threadIdx.x == 0) iters[0] = it;
        } else if (threadIdx.x < 64) {
            const int lane = threadIdx.x;
            int it = 0;
            while (true) {
                wv_fill_ghosts(L, UU(q), D.bc, lane);
                wv_sweep(L, UU(q), FF(q), CX(q), CY(q), CZ(q), 0, lane);
                wv_fill_ghosts(L, UU(q), D.bc, lane);
                wv_sweep(L, UU(q), FF(q), CX(q), CY(q), CZ(q), 1, lane);
                double gm = wv_max_res<T>(L, UU(q), FF(q), CX(q), CY(q), CZ(q), lane);
                if (gm < D.tol) break;
                it += 1;
                if (it > D.max_count) break;
            }
            wv_fill_ghosts(L, UU(q), D.bc, lane);
            if (lane == 0) iters[0] = it;
        }
        __syncthreads();
        }
    }
    // up leg
    const bool images = D.bc.rule[0] <= G_ANTI && D.bc.rule[1] <= G_ANTI && D.bc.rule[2] <= G_ANTI;
    for (int q = nl - 2; q >= 0; q--) {
        const Lay &L = D.L[q];
        blk_prolong_add<T>(L, UU(q), D.L[q + 1], UU(q + 1), D.bc, images, D.post > 0 ? 1 : -1);
        if (!images) lds_fill_ghosts(L, UU(q), D.bc);
        for (int s = 0; s < D.post; s++) {
            blk_sweep<T>(L, UU(q), FF(q), CX(q), CY(q), CZ(q), 0, D.bc, false);
            blk_sweep<T>(L, UU(q), FF(q), CX(q), CY(q), CZ(q), 1, D.bc, false);
        }
    }
    {
        long n = lay_size(D.L[0]);
        T *u = UU(0);
        for (long t = threadIdx.x; t < n; t += blockDim.x) u_first[t] = u[t];
    }
#undef CX
#undef CY
#undef CZ
#undef UU
#undef FF
}

// ------------------------------------------------------------------------------------------------
// launch wrappers
// ------------------------------------------------------------------------------------------------
static inline unsigned nblk(long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

static Geo geo_planes(unsigned per_plane, unsigned ip0, unsigned nplanes, unsigned d1) {
    Geo g;
    g.per_plane = per_plane;
    g.bpp = (per_plane + 255u) / 256u;
    g.ip0 = ip0;
    g.nplanes = nplanes;
    g.total_blocks = g.bpp * nplanes;
    g.d1 = d1;
    g.inv1 = 1.0f / (float)d1;
    return g;
}
// one colour of the core nodes: planes 1..nx, slots (jp, h) with jp in 1..ny
static Geo geo_sweep(const Lay &L) { return geo_planes((unsigned)(L.ny * L.kh), 1, L.nx, L.kh); }
// both colours of the core nodes: slots (jp, c, h)
static Geo geo_core(const Lay &L) { return geo_planes((unsigned)(L.ny * 2 * L.kh), 1, L.nx, 2 * L.kh); }
// both colours of the core nodes, four x-planes per thread (core4_visit)
static Geo geo_core4(const Lay &L) { return geo_planes((unsigned)(L.ny * 2 * L.kh), 0, (L.nx + 3) / 4, 2 * L.kh); }
// one colour of geo_core4
static Geo geo_core4c(const Lay &L) { return geo_planes((unsigned)(L.ny * L.kh), 0, (L.nx + 3) / 4, L.kh); }
// every stored slot of the padded level
static Geo geo_all(const Lay &L) { return geo_planes((unsigned)((L.ny + 2) * 2 * L.khs), 0, L.nx + 2, 2 * L.khs); }
static const unsigned REDUCE_BLOCKS = 2048;
// a multiple of 8 (one share per XCD), at most REDUCE_BLOCKS
static unsigned reduce_grid(const Geo &g) {
    unsigned n = g.total_blocks < REDUCE_BLOCKS ? g.total_blocks : REDUCE_BLOCKS;
    return (n + 7u) & ~7u;
}

template <typename T>
void launch_fill_ghosts(const Lay &L, T *a, const Bc &bc, hipStream_t s) {
    long total = 2L * (L.nx + 2) * (L.ny + 2) + 2L * (L.nx + 2) * L.nz + 2L * L.ny * L.nz;
    hipLaunchKernelGGL(k_fill_ghosts<T>, dim3(nblk(total)), dim3(256), 0, s, L, a, bc, total);
}

template <typename T>
void launch_rb_sweep(const Lay &L, T *u, const T *f, const T *cx, const T *cy, const T *cz, int color,
                     const Bc &bc, bool zero_start, hipStream_t s) {
    Geo G = geo_sweep(L);
    if (zero_start)
        hipLaunchKernelGGL((k_rb_sweep<T, true>), dim3(G.total_blocks), dim3(256), 0, s, L, G, u, f, cx, cy, cz, color, bc);
    else
        hipLaunchKernelGGL((k_rb_sweep<T, false>), dim3(G.total_blocks), dim3(256), 0, s, L, G, u, f, cx, cy, cz, color, bc);
}

template <typename T, typename TO>
void launch_residual(const Lay &L, const T *u, const T *f, TO *t, const T *cx, const T *cy, const T *cz,
                     hipStream_t s, int color, const Bc *img) {
    Geo G = color >= 0 ? geo_sweep(L) : geo_core(L);
    if (img)
        hipLaunchKernelGGL((k_residual<T, TO, true>), dim3(G.total_blocks), dim3(256), 0, s, L, G, u, f, t, cx, cy,
                           cz, color, *img);
    else
        hipLaunchKernelGGL((k_residual<T, TO, false>), dim3(G.total_blocks), dim3(256), 0, s, L, G, u, f, t, cx, cy,
                           cz, color, Bc{});
}

template <typename T>
void launch_restrict(const Lay &F, const Lay &Cl, const T *t, T *fc, hipStream_t s, bool red0,
                     const ZeroStart<T> *zs) {
    Geo G = geo_core(Cl);
    const ZeroStart<T> z = zs ? *zs : ZeroStart<T>{};
    if (red0 && zs)
        hipLaunchKernelGGL((k_restrict<T, true, true>), dim3(G.total_blocks), dim3(256), 0, s, F, Cl, G, t, fc, z);
    else if (red0)
        hipLaunchKernelGGL((k_restrict<T, true, false>), dim3(G.total_blocks), dim3(256), 0, s, F, Cl, G, t, fc, z);
    else if (zs)
        hipLaunchKernelGGL((k_restrict<T, false, true>), dim3(G.total_blocks), dim3(256), 0, s, F, Cl, G, t, fc, z);
    else
        hipLaunchKernelGGL((k_restrict<T, false, false>), dim3(G.total_blocks), dim3(256), 0, s, F, Cl, G, t, fc, z);
}

template <typename T, typename TF>
void launch_prolong(const Lay &F, const Lay &Cl, const T *e, TF *u, int accumulate, hipStream_t s, int color,
                    const Bc *img) {
    Geo G = color >= 0 ? geo_sweep(F) : geo_core(F);
    if (img)
        hipLaunchKernelGGL((k_prolong<T, TF, true>), dim3(G.total_blocks), dim3(256), 0, s, F, Cl, G, e, u,
                           accumulate, color, *img);
    else
        hipLaunchKernelGGL((k_prolong<T, TF, false>), dim3(G.total_blocks), dim3(256), 0, s, F, Cl, G, e, u,
                           accumulate, color, Bc{});
}

template <typename T>
void launch_fill(T *a, long n, T v, hipStream_t s) {
    unsigned b = nblk(n);
    if (b > 65536u) b = 65536u;
    if (b == 0) b = 1;
    hipLaunchKernelGGL(k_fill<T>, dim3(b), dim3(256), 0, s, a, n, v);
}

template <typename T>
void launch_add_scalar(const Lay &L, T *a, const double *dv, double scale, int core, hipStream_t s) {
    Geo G = geo_all(L);
    hipLaunchKernelGGL(k_add_scalar<T>, dim3(G.total_blocks), dim3(256), 0, s, L, G, a, dv, scale, core);
}

template <typename TS, typename TD>
void launch_pack(const Lay &L, const TS *ref, TD *dst, int core_only, hipStream_t s) {
    long n = (long)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    hipLaunchKernelGGL((k_pack<TS, TD>), dim3(nblk(n)), dim3(256), 0, s, L, ref, dst, core_only);
}

template <typename TS, typename TD>
void launch_pack_full(const Lay &L, const TS *ref, TD *dst, hipStream_t s) {
    hipLaunchKernelGGL((k_pack_full<TS, TD>), dim3(nblk(lay_size(L))), dim3(256), 0, s, L, ref, dst);
}

template <typename TS, typename TD>
void launch_unpack(const Lay &L, const TS *src, TD *ref, hipStream_t s) {
    long n = (long)(L.nx + 2) * (L.ny + 2) * (L.nz + 2);
    hipLaunchKernelGGL((k_unpack<TS, TD>), dim3(nblk(n)), dim3(256), 0, s, L, src, ref);
}

int norm_parts(const Lay &L) { return (int)reduce_grid(geo_all(L)); }

template <typename T>
void launch_rhs_sums(const Lay &L, const T *f, const double *wx, const double *wy, const double *wz, double *part,
                     double *out, hipStream_t s) {
    Geo G = geo_all(L);
    int nb = (int)reduce_grid(G);
    hipLaunchKernelGGL(k_rhs_sums<T>, dim3(nb), dim3(256), 0, s, L, G, f, wx, wy, wz, part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s, part, nb, out);
}

template <typename T>
size_t coarsest_lds_bytes(const Lay &L) {
    return sizeof(T) * (2 * (size_t)lay_size(L) + 4 * (size_t)(L.nx + L.ny + L.nz + 6));
}

template <typename T>
void launch_coarsest(const Lay &L, T *u, const T *f, const T *cx, const T *cy, const T *cz, const Bc &bc,
                     double tol, int max_count, int *iters, hipStream_t s) {
    size_t lds = coarsest_lds_bytes<T>(L);
    hipLaunchKernelGGL(k_coarsest<T>, dim3(1), dim3(256), lds, s, L, u, f, cx, cy, cz, bc, tol, max_count, iters);
}

template <typename T>
void launch_jacobi(const Lay &L, const T *u, const T *f, T *t, const T *cx, const T *cy, const T *cz,
                   hipStream_t s) {
    Geo G = geo_core(L);
    hipLaunchKernelGGL(k_jacobi<T>, dim3(G.total_blocks), dim3(256), 0, s, L, G, u, f, t, cx, cy, cz);
}

template <typename T>
void launch_project_small(const Lay &L, T *f, const double *wx, const double *wy, const double *wz, double wsum,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_project_small<T>, dim3(1), dim3(256), 0, s, L, f, wx, wy, wz, wsum);
}

void launch_outer_residual(const Lay &L, const double *u_in, const float *e, const double *f, double *u_out,
                           float *r32, const double *cx, const double *cy, const double *cz, const Bc &bc,
                           double *part2, const double *fterms, double *out, hipStream_t s) {
    Geo G = geo_core4(L);
    int nb = (int)G.total_blocks;
    if (e)
        hipLaunchKernelGGL(k_outer_residual<true>, dim3(nb), dim3(256), 0, s, L, G, u_in, e, f, u_out, r32, cx, cy,
                           cz, bc, part2);
    else
        hipLaunchKernelGGL(k_outer_residual<false>, dim3(nb), dim3(256), 0, s, L, G, u_in, e, f, u_out, r32, cx, cy,
                           cz, bc, part2);
    hipLaunchKernelGGL(k_norm_combine, dim3(1), dim3(1024), 0, s, part2, nb, fterms, out);
}

void launch_gcr_dots(const Lay &L, const float *e, const float *z, const float *r32, const float *cx,
                     const float *cy, const float *cz, double *part5, double *coef, bool q1r, hipStream_t s,
                     const FinalBufs *fb) {
    Geo G = geo_core4(L);
    int nb = (int)G.total_blocks;
    if (q1r)
        hipLaunchKernelGGL(k_gcr_dots<true>, dim3(nb), dim3(256), 0, s, L, G, e, z, r32, cx, cy, cz, part5);
    else
        hipLaunchKernelGGL(k_gcr_dots<false>, dim3(nb), dim3(256), 0, s, L, G, e, z, r32, cx, cy, cz, part5);
    if (fb)
        hipLaunchKernelGGL(k_gcr_coef_mb, dim3(FINAL_BLOCKS), dim3(256), 0, s, part5, nb, coef, fb->stage,
                           fb->counter);
    else
        hipLaunchKernelGGL(k_gcr_coef, dim3(1), dim3(1024), 0, s, part5, nb, coef);
}

void launch_gcr_step(const Lay &L, const double *x_in, const float *e, const float *z, const double *coef,
                     double *x_out, float *z_next, const Bc &bc, const Bc &ebc, hipStream_t s) {
    Geo G = geo_core4(L);
    hipLaunchKernelGGL(k_gcr_step, dim3(G.total_blocks), dim3(256), 0, s, L, G, x_in, e, z, coef, x_out, z_next, bc,
                       ebc);
}

int norm_res_parts(const Lay &L) { return (int)geo_core4(L).total_blocks; }

template <typename T>
void launch_norm_res(const Lay &L, const T *u, const T *f, const T *cx, const T *cy, const T *cz, double *part2,
                     const double *fterms, double *out, hipStream_t s, int color) {
    Geo G = color >= 0 ? geo_core4c(L) : geo_core4(L);
    int nb = (int)G.total_blocks;
    hipLaunchKernelGGL(k_norm_res<T>, dim3(nb), dim3(256), 0, s, L, G, u, f, cx, cy, cz, part2, color);
    hipLaunchKernelGGL(k_norm_combine, dim3(1), dim3(1024), 0, s, part2, nb, fterms, out);
}

template <typename T>
void launch_f_terms(const Lay &L, const T *f, double *part, double *fterms, hipStream_t s) {
    Geo G = geo_all(L);
    int nb = (int)reduce_grid(G);
    hipLaunchKernelGGL(k_f_terms<T>, dim3(nb), dim3(256), 0, s, L, G, f, part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s, part, nb, fterms);
}

void launch_wsum_ref(const Lay &L, const double *ref, const double *wx, const double *wy, const double *wz,
                     double *part, double *out, hipStream_t s) {
    // in-plane entries below 2^24 (fdivu); at most REDUCE_BLOCKS partials (the rows of part)
    const unsigned n2 = (unsigned)L.nz + 2;
    const unsigned bpp = ((unsigned)(L.ny + 2) * n2 + 255u) / 256u;
    unsigned groups = REDUCE_BLOCKS / bpp;
    if (groups > (unsigned)L.nx) groups = (unsigned)L.nx;
    if (groups < 1u) groups = 1u;
    const unsigned nb = bpp * groups < REDUCE_BLOCKS ? bpp * groups : REDUCE_BLOCKS;
    hipLaunchKernelGGL(k_wsum_ref, dim3(nb), dim3(256), 0, s, L, bpp, groups, 1.0f / (float)n2, ref, wx, wy, wz,
                       part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s, part, (int)nb, out);
}

void launch_pack_shift(const Lay &L, const double *ref, double *dst, const double *q, double scale, double *part,
                       double *fterms, hipStream_t s) {
    Geo G = geo_all(L);
    int nb = (int)reduce_grid(G);
    hipLaunchKernelGGL(k_pack_shift, dim3(nb), dim3(256), 0, s, L, G, ref, dst, q, scale, part);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, s, part, nb, fterms);
}

template <typename T>
__global__ void __launch_bounds__(1024) k_coarsest_direct(Lay L, T *__restrict__ u, const T *__restrict__ f,
                                                          const T *__restrict__ pt, int nc, Bc bc,
                                                          int *__restrict__ iters) {
    long n = lay_size(L);
    for (long t = threadIdx.x; t < n; t += blockDim.x) u[t] = T(0);
    __syncthreads();
    blk_direct<T>(L, u, f, pt, nc);
    lds_fill_ghosts(L, u, bc);
    if (threadIdx.x == 0) iters[0] = 0;
}

template <typename T>
void launch_coarsest_direct(const Lay &L, T *u, const T *f, const T *pt, int nc, const Bc &bc, int *iters,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_coarsest_direct<T>, dim3(1), dim3(1024), 0, s, L, u, f, pt, nc, bc, iters);
}

template <typename T>
size_t coarse_cycle_plan(CoarseDesc &D, size_t limit) {
    size_t off = 0;
    auto take = [&](size_t n) { size_t o = off; off += (n + 7) & ~(size_t)7; return (int)o; };
    for (int q = 0; q <= D.last - D.first; q++) {
        const Lay &L = D.L[q];
        D.off_u[q] = take((size_t)lay_size(L));
        D.off_f[q] = take((size_t)lay_size(L));
        D.off_c[q][0] = take(4 * (size_t)(L.nx + 2));
        D.off_c[q][1] = take(4 * (size_t)(L.ny + 2));
        D.off_c[q][2] = take(4 * (size_t)(L.nz + 2));
    }
    // scratch of the tap-parallel restriction, where it fits
    D.off_s = -1;
    D.tap_max = 4 * kTapNodes;
    if ((off + kTapScratch + 8) * sizeof(T) <= limit) D.off_s = take(kTapScratch);
    return off * sizeof(T);
}

template <typename T>
void launch_coarse_cycle(const CoarseDesc &D, size_t lds, T *u_first, const T *f_first, int *iters, hipStream_t s) {
    hipLaunchKernelGGL(k_coarse_cycle<T>, dim3(1), dim3(1024), lds, s, D, u_first, f_first, iters);
}

template <typename T>
int coarse_cycle_attr(size_t lds) {
    return hipFuncSetAttribute((const void *)k_coarse_cycle<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) ==
                   hipSuccess ? 0 : 1;
}

// explicit instantiations
#define SR_INST(T)                                                                                          \
    template void launch_fill_ghosts<T>(const Lay &, T *, const Bc &, hipStream_t);                        \
    template void launch_rb_sweep<T>(const Lay &, T *, const T *, const T *, const T *, const T *, int,     \
                                     const Bc &, bool, hipStream_t);                                        \
    template void launch_restrict<T>(const Lay &, const Lay &, const T *, T *, hipStream_t, bool,          \
                                     const ZeroStart<T> *);                                                 \
    template void launch_fill<T>(T *, long, T, hipStream_t);                                               \
    template void launch_add_scalar<T>(const Lay &, T *, const double *, double, int, hipStream_t);        \
    template void launch_norm_res<T>(const Lay &, const T *, const T *, const T *, const T *, const T *,   \
                                     double *, const double *, double *, hipStream_t, int);                 \
    template void launch_f_terms<T>(const Lay &, const T *, double *, double *, hipStream_t);               \
    template void launch_coarsest_direct<T>(const Lay &, T *, const T *, const T *, int, const Bc &, int *,    \
                                            hipStream_t);                                                  \
    template void launch_rhs_sums<T>(const Lay &, const T *, const double *, const double *, const double *, \
                                     double *, double *, hipStream_t);                                      \
    template size_t coarsest_lds_bytes<T>(const Lay &);                                                     \
    template void launch_coarsest<T>(const Lay &, T *, const T *, const T *, const T *, const T *,         \
                                     const Bc &, double, int, int *, hipStream_t);            \
    template void launch_jacobi<T>(const Lay &, const T *, const T *, T *, const T *, const T *, const T *,  \
                                   hipStream_t);                                                            \
    template void launch_project_small<T>(const Lay &, T *, const double *, const double *, const double *, \
                                          double, hipStream_t);
SR_INST(double)
SR_INST(float)
template size_t coarse_cycle_plan<double>(CoarseDesc &, size_t);
template size_t coarse_cycle_plan<float>(CoarseDesc &, size_t);
template void launch_coarse_cycle<double>(const CoarseDesc &, size_t, double *, const double *, int *, hipStream_t);
template void launch_coarse_cycle<float>(const CoarseDesc &, size_t, float *, const float *, int *, hipStream_t);
template int coarse_cycle_attr<double>(size_t);
template int coarse_cycle_attr<float>(size_t);
#undef SR_INST

template void launch_residual<double, double>(const Lay &, const double *, const double *, double *,
                                              const double *, const double *, const double *, hipStream_t, int, const Bc *);
template void launch_residual<float, float>(const Lay &, const float *, const float *, float *, const float *,
                                            const float *, const float *, hipStream_t, int, const Bc *);
template void launch_residual<double, float>(const Lay &, const double *, const double *, float *,
                                             const double *, const double *, const double *, hipStream_t, int, const Bc *);
template void launch_prolong<double, double>(const Lay &, const Lay &, const double *, double *, int, hipStream_t, int, const Bc *);
template void launch_prolong<float, float>(const Lay &, const Lay &, const float *, float *, int, hipStream_t, int, const Bc *);
template void launch_prolong<float, double>(const Lay &, const Lay &, const float *, double *, int, hipStream_t, int, const Bc *);
template void launch_pack<double, double>(const Lay &, const double *, double *, int, hipStream_t);
template void launch_pack<double, float>(const Lay &, const double *, float *, int, hipStream_t);
template void launch_pack_full<double, double>(const Lay &, const double *, double *, hipStream_t);
template void launch_unpack<double, double>(const Lay &, const double *, double *, hipStream_t);
template void launch_unpack<float, double>(const Lay &, const float *, double *, hipStream_t);

}  // namespace sr
