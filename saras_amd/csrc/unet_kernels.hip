// gfx950 kernels of the learned initial guess: implicit-GEMM 3x3x3 convolutions on the matrix
// cores (v_mfma_f32_16x16x32_bf16: bf16 operands, fp32 accumulation), their weight gradients,
// the transfer operators between the network's node-centred levels, and the optimizer.
//
// Tensors are channels-last, voxel v = (x * ny + y) * nz + z, C channels contiguous, bf16.
// Convolution as a GEMM per output tile: D[co][voxel] += W[co][k] X[k][voxel], with the
// reduction index k = (tap, input channel) and the input voxel = output voxel + tap offset
// (zero outside the grid).  A workgroup owns a 4 x 4 x 16 output tile; its 6 x 6 x 18 input
// halo is staged once per 32-channel chunk in LDS; each of its four waves computes the four
// z-lines of one x plane for every block of 16 output channels.
#include <hip/hip_runtime.h>

#include "unet_kernels.h"

namespace sr {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

// round to nearest, ties to even (finite inputs)
__device__ __forceinline__ bf16_t f2bf(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ bf16x8 zero_frag() {
    u32x4 z = {0u, 0u, 0u, 0u};
    return __builtin_bit_cast(bf16x8, z);
}

template <int CIN>
struct ConvShape {
    static constexpr int CH = CIN < 32 ? CIN : 32;   // channels staged per chunk
    static constexpr int NCH = CIN / CH;               // chunks
    static constexpr int NS = (27 * CH + 31) / 32;     // K-steps of 32 per chunk
    // LDS voxel stride (bf16 elements): an odd number of 16-byte units keeps the 16 lanes of a
    // fragment read on distinct banks
    static constexpr int VS = ((CH * 2 / 16) % 2 == 0) ? CH + 8 : CH;
};

static int conv_ns(int cin) {
    int ch = cin < 32 ? cin : 32;
    return (27 * ch + 31) / 32;
}

long conv_pack_elems(int cin, int cout) {
    int ch = cin < 32 ? cin : 32;
    return (long)(cin / ch) * conv_ns(cin) * (cout / 16) * 64 * 8;
}

// packed[(((c * NS + s) * NCB + cb) * 64 + lane) * 8 + i] = W'[tap][co][ci] with
// co = cb * 16 + lane % 16, kk = s * 32 + 8 (lane / 16) + i, tap = kk / CH, ci = c * CH + kk % CH
__global__ void k_pack(const float *__restrict__ W, int cin_l, int cout_l, int transpose, int cin, int cout,
                       long total, bf16_t *__restrict__ packed) {
    long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    int ch = cin < 32 ? cin : 32;
    int ns = (27 * ch + 31) / 32, ncb = cout / 16;
    int i = (int)(e % 8);
    long r = e / 8;
    int lane = (int)(r % 64);
    r /= 64;
    int cb = (int)(r % ncb);
    r /= ncb;
    int s = (int)(r % ns);
    int c = (int)(r / ns);
    int co = cb * 16 + (lane & 15);
    int kk = s * 32 + 8 * (lane >> 4) + i;
    int tap = kk / ch, ci = c * ch + kk % ch;
    float v = 0.0f;
    if (tap < 27) {
        // forward: W[tap][co][ci]; adjoint: W[26 - tap][ci][co] (mirrored taps, swapped channels)
        v = transpose ? W[((long)(26 - tap) * cout_l + ci) * cin_l + co] : W[((long)tap * cout_l + co) * cin_l + ci];
    }
    packed[e] = f2bf(v);
}

template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_conv(NetGrid g, const bf16_t *__restrict__ in,
                                              const bf16_t *__restrict__ wpack, const float *__restrict__ bias,
                                              void *__restrict__ out, int mode) {
    using K = ConvShape<CIN>;
    constexpr int CH = K::CH, NCH = K::NCH, NS = K::NS, VS = K::VS, NCB = COUT / 16, PARTS = CH / 8;
    __shared__ __attribute__((aligned(16))) bf16_t halo[648 * VS];
    const int tile = blockIdx.x;
    const int tzi = tile % g.tz, rr = tile / g.tz, tyi = rr % g.ty, txi = rr / g.ty;
    const int x0 = txi * 4, y0 = tyi * 4, z0 = tzi * 16;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m = lane & 15, gq = lane >> 4;
    f32x4 acc[4][NCB];
#pragma unroll
    for (int yy = 0; yy < 4; yy++)
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) acc[yy][cb] = (f32x4){0.f, 0.f, 0.f, 0.f};

    for (int c = 0; c < NCH; c++) {
        if (c) __syncthreads();
        for (int idx = threadIdx.x; idx < 648 * PARTS; idx += 256) {
            int v = idx / PARTS, p = idx % PARTS;
            int hz = v % 18, hy = (v / 18) % 6, hx = v / 108;
            int x = x0 + hx - 1, y = y0 + hy - 1, z = z0 + hz - 1;
            u32x4 val = {0u, 0u, 0u, 0u};
            if (x >= 0 && x < g.nx && y >= 0 && y < g.ny && z >= 0 && z < g.nz)
                val = *(const u32x4 *)(in + (((long)x * g.ny + y) * g.nz + z) * CIN + c * CH + p * 8);
            *(u32x4 *)(halo + v * VS + p * 8) = val;
        }
        __syncthreads();
        for (int s = 0; s < NS; s++) {
            const int kk = s * 32 + 8 * gq;
            const int tap = kk / CH, cc = kk % CH;
            const bool valid = tap < 27;
            const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
            bf16x8 a[NCB];
#pragma unroll
            for (int cb = 0; cb < NCB; cb++)
                a[cb] = *(const bf16x8 *)(wpack + ((((long)c * NS + s) * NCB + cb) * 64 + lane) * 8);
#pragma unroll
            for (int yy = 0; yy < 4; yy++) {
                bf16x8 b = zero_frag();
                if (valid) {
                    int hv = ((w + 1 + dx) * 6 + (yy + 1 + dy)) * 18 + (m + 1 + dz);
                    b = *(const bf16x8 *)(halo + hv * VS + cc);
                }
#pragma unroll
                for (int cb = 0; cb < NCB; cb++)
                    acc[yy][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[cb], b, acc[yy][cb], 0, 0, 0);
            }
        }
    }
    // lane holds D[co = cb * 16 + 4 gq + r][voxel z0 + m] of each line
    const int x = x0 + w, z = z0 + m;
#pragma unroll
    for (int yy = 0; yy < 4; yy++) {
        const int y = y0 + yy;
        if (x >= g.nx || y >= g.ny || z >= g.nz) continue;
        const long v = ((long)x * g.ny + y) * g.nz + z;
#pragma unroll
        for (int cb = 0; cb < NCB; cb++) {
            const int co0 = cb * 16 + 4 * gq;
            if (mode == CONV_BF16_RELU) {
                float r0 = fmaxf(acc[yy][cb][0] + bias[co0 + 0], 0.f);
                float r1 = fmaxf(acc[yy][cb][1] + bias[co0 + 1], 0.f);
                float r2 = fmaxf(acc[yy][cb][2] + bias[co0 + 2], 0.f);
                float r3 = fmaxf(acc[yy][cb][3] + bias[co0 + 3], 0.f);
                u32x2 pk = {(unsigned)f2bf(r0) | ((unsigned)f2bf(r1) << 16),
                            (unsigned)f2bf(r2) | ((unsigned)f2bf(r3) << 16)};
                *(u32x2 *)((bf16_t *)out + v * COUT + co0) = pk;
            } else if (mode == CONV_F32) {
                *(f32x4 *)((float *)out + v * COUT + co0) = acc[yy][cb];
            } else if (cb == 0 && gq == 0) {
                ((float *)out)[v] = acc[yy][0][0] + bias[0];
            }
        }
    }
}

// Weight gradient of one (16-channel output block, 16-channel input block) pair:
// D[co][ci] += dY^T[co][voxel] X[voxel + tap][ci], summed over the tiles of one slot.  The
// contraction runs over voxels, which are strided in the channels-last tensors, so both tiles are
// staged in LDS transposed -- channel-major with the z voxels contiguous -- and the input halo as
// three copies shifted by dz = -1, 0, 1: every MFMA operand (8 consecutive z voxels of one channel)
// is then one aligned 16-byte LDS read.  Wave w owns the taps w, w + 4, ..., so each tap's
// accumulator lives in exactly one wave; the A fragment (output gradient) of a 32-voxel K-step is
// shared by the wave's taps.  Row strides: 264 and 584 elements (132 and 292 dwords), so the 16
// lanes of a fragment read hit disjoint banks.
constexpr int WG_GST = 264;   // gT row: 256 voxels (16 lines x 16 z) + 8
constexpr int WG_XST = 584;   // xT channel: 36 halo lines (6 x 6) x 16 z + 8
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_wgrad(NetGrid g, const bf16_t *__restrict__ dy, const bf16_t *__restrict__ x,
                                               float *__restrict__ part, int nslots, float *__restrict__ bpart) {
    constexpr int NIB = (CIN + 15) / 16, NPAIR = (COUT / 16) * NIB;
    const int pair = blockIdx.y, cb = pair / NIB, ib = pair % NIB;
    const int slot = blockIdx.x;
    __shared__ __attribute__((aligned(16))) bf16_t gT[16 * WG_GST];
    __shared__ __attribute__((aligned(16))) bf16_t xT[3 * 16 * WG_XST];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int m = lane & 15, gq = lane >> 4;
    f32x4 acc[7];
#pragma unroll
    for (int j = 0; j < 7; j++) acc[j] = (f32x4){0.f, 0.f, 0.f, 0.f};
    const long ntiles = g.ntiles();
    // The next tile's halo and gradient items are loaded into registers while the current tile
    // computes (each thread owns items tid + 256 k: at most 6 halo items and one gradient pair).
    u32x4 xr[6], ga, gb;
    auto load = [&](long tile) {
        const int tzi = (int)(tile % g.tz), rr = (int)(tile / g.tz), tyi = rr % g.ty, txi = rr / g.ty;
        const int x0 = txi * 4, y0 = tyi * 4, z0 = tzi * 16;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const int idx = threadIdx.x + 256 * k;
            xr[k] = (u32x4){0u, 0u, 0u, 0u};
            if (idx < 648 * 2) {
                const int v = idx >> 1, p = idx & 1;
                const int hz = v % 18, hl = v / 18, hy = hl % 6, hx = hl / 6;
                const int xx = x0 + hx - 1, yv = y0 + hy - 1, zz = z0 + hz - 1;
                if (xx >= 0 && xx < g.nx && yv >= 0 && yv < g.ny && zz >= 0 && zz < g.nz && ib * 16 + p * 8 < CIN)
                    xr[k] = *(const u32x4 *)(x + (((long)xx * g.ny + yv) * g.nz + zz) * CIN + ib * 16 + p * 8);
            }
        }
        {
            const int idx = threadIdx.x;
            const int vp = idx >> 1, p = idx & 1;
            const int lz = (vp % 8) * 2, L = vp / 8, ly = L & 3, lx = L >> 2;
            const int xx = x0 + lx, yv = y0 + ly, zz = z0 + lz;
            ga = (u32x4){0u, 0u, 0u, 0u};
            gb = (u32x4){0u, 0u, 0u, 0u};
            if (xx < g.nx && yv < g.ny) {
                const long base = (((long)xx * g.ny + yv) * g.nz + zz) * COUT + cb * 16 + p * 8;
                if (zz < g.nz) ga = *(const u32x4 *)(dy + base);
                if (zz + 1 < g.nz) gb = *(const u32x4 *)(dy + base + COUT);
            }
        }
    };
    // input halo (6 x 6 x 18 voxels, 8 channels per item) into the three shifted copies: copy c
    // holds halo z = z' + c at z' = 0..15; the gradient tile (4 x 4 x 16 voxels) as (z, z + 1) pairs
    auto store = [&]() {
        unsigned short *xs16 = (unsigned short *)xT;
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const int idx = threadIdx.x + 256 * k;
            if (idx >= 648 * 2) continue;
            const int v = idx >> 1, p = idx & 1;
            const int hz = v % 18, hl = v / 18;
            const u32x4 val = xr[k];
            const unsigned short e[8] = {(unsigned short)(val[0] & 0xffffu), (unsigned short)(val[0] >> 16),
                                         (unsigned short)(val[1] & 0xffffu), (unsigned short)(val[1] >> 16),
                                         (unsigned short)(val[2] & 0xffffu), (unsigned short)(val[2] >> 16),
                                         (unsigned short)(val[3] & 0xffffu), (unsigned short)(val[3] >> 16)};
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const int zp = hz - c;
                if (zp < 0 || zp > 15) continue;
#pragma unroll
                for (int j = 0; j < 8; j++) xs16[(c * 16 + p * 8 + j) * WG_XST + hl * 16 + zp] = e[j];
            }
        }
        {
            const int idx = threadIdx.x;
            const int vp = idx >> 1, p = idx & 1;
            const int lz = (vp % 8) * 2, L = vp / 8;
            unsigned int *gs32 = (unsigned int *)gT;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                // channels 2j and 2j + 1: (z, z + 1) pairs
                const unsigned lo0 = ga[j] & 0xffffu, hi0 = ga[j] >> 16, lo1 = gb[j] & 0xffffu, hi1 = gb[j] >> 16;
                gs32[((p * 8 + 2 * j) * WG_GST + L * 16 + lz) >> 1] = lo0 | (lo1 << 16);
                gs32[((p * 8 + 2 * j + 1) * WG_GST + L * 16 + lz) >> 1] = hi0 | (hi1 << 16);
            }
        }
    };
    // bias gradient of the output block: the blocks of input block 0 also add the staged gradient
    // tile per channel (channel threadIdx / 16, voxels threadIdx % 16 + 16 k, then the 16 lanes)
    const bool bias = bpart != nullptr && ib == 0;
    const int bco = threadIdx.x >> 4, bsub = threadIdx.x & 15;
    float bacc = 0.f;
    if (slot < ntiles) load(slot);
    for (long tile = slot; tile < ntiles; tile += nslots) {
        __syncthreads();
        store();
        __syncthreads();
        if (tile + nslots < ntiles) load(tile + nslots);
        if (bias) {
            const unsigned short *g16 = (const unsigned short *)gT;
            float t = 0.f;
#pragma unroll
            for (int k = 0; k < 16; k++) t += __uint_as_float((unsigned)g16[bco * WG_GST + bsub + 16 * k] << 16);
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
            bacc += t;
        }
#pragma unroll 2
        for (int q = 0; q < 8; q++) {
            // K-step q: voxels 8 gq + i (i = 0..7) = z 8 (gq & 1) + i of line 2 q + (gq >> 1)
            const int L = 2 * q + (gq >> 1), zo = 8 * (gq & 1);
            const bf16x8 A = *(const bf16x8 *)(gT + m * WG_GST + L * 16 + zo);
            const int lx = L >> 2, ly = L & 3;
#pragma unroll
            for (int j = 0; j < 7; j++) {
                const int tap = w + 4 * j;
                if (tap < 27) {
                    const int dx = tap / 9 - 1, dyy = (tap / 3) % 3 - 1, c = tap % 3;
                    const int hl = (lx + 1 + dx) * 6 + (ly + 1 + dyy);
                    const bf16x8 B = *(const bf16x8 *)(xT + (c * 16 + m) * WG_XST + hl * 16 + zo);
                    acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc[j], 0, 0, 0);
                }
            }
        }
    }
    if (bias && bsub == 0) bpart[(long)slot * COUT + cb * 16 + bco] = bacc;
    // D[co = 4 gq + r][ci = m] of tap w + 4 j
#pragma unroll
    for (int j = 0; j < 7; j++) {
        const int tap = w + 4 * j;
        if (tap < 27) {
            float *dst = part + (((long)slot * NPAIR + pair) * 27 + tap) * 256;
#pragma unroll
            for (int r = 0; r < 4; r++) dst[(4 * gq + r) * 16 + m] = acc[j][r];
        }
    }
}

// dW[tap][co][ci] = sum over slots of the partials.  A block reduces 16 consecutive outputs (one
// channel pair and tap, consecutive ci): 16 threads per output each add the slots j, j + 16, ... in
// order, then one thread per output adds the 16 partial sums in order -- a fixed order, so the
// result is the same on every run.
__global__ void __launch_bounds__(256) k_wgrad_final(const float *__restrict__ part, int cin, int cout, int nslots,
                                                     float *__restrict__ dW) {
    const long total = 27L * cout * cin;
    const int o = threadIdx.x & 15, j = threadIdx.x >> 4;
    const long e = (long)blockIdx.x * 16 + o;
    float s = 0.f;
    if (e < total) {
        int ci = (int)(e % cin), co = (int)((e / cin) % cout), tap = (int)(e / ((long)cin * cout));
        int nib = (cin + 15) / 16, npair = (cout / 16) * nib;
        int pair = (co / 16) * nib + ci / 16;
        int idx = (co % 16) * 16 + (ci % 16);
        const float *p = part + ((long)pair * 27 + tap) * 256 + idx;
        const long stride = (long)npair * 27 * 256;
        for (int sl = j; sl < nslots; sl += 16) s += p[(long)sl * stride];
    }
    __shared__ float red[16][17];
    red[j][o] = s;
    __syncthreads();
    if (threadIdx.x < 16 && e < total) {
        float t = 0.f;
        for (int q = 0; q < 16; q++) t += red[q][threadIdx.x];
        dW[e] = t;
    }
}

// db[ch] = sum over slots of the partials: 256 / cout threads per channel (cout divides 256) each
// add the slots lane, lane + 256 / cout, ... in order, then one thread per channel adds those sums in
// order (a fixed order, the same result on every run)
__global__ void __launch_bounds__(256) k_bias_final(const float *__restrict__ part, int cout, int nslots,
                                                    float *__restrict__ db) {
    const int per = 256 / cout;
    const int ch = threadIdx.x % cout, lane = threadIdx.x / cout;
    float s = 0.f;
    if (lane < per)
        for (int sl = lane; sl < nslots; sl += per) s += part[(long)sl * cout + ch];
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < cout) {
        float t = 0.f;
        for (int q = 0; q < per; q++) t += red[q * cout + threadIdx.x];
        db[threadIdx.x] = t;
    }
}

// ------------------------------------------------------------------------------------------------
// transfer operators (one thread per voxel and 4 channels)
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void load4(const bf16_t *p, float *v) {
    u32x2 u = *(const u32x2 *)p;
    v[0] = __uint_as_float(u[0] << 16);
    v[1] = __uint_as_float(u[0] & 0xFFFF0000u);
    v[2] = __uint_as_float(u[1] << 16);
    v[3] = __uint_as_float(u[1] & 0xFFFF0000u);
}
__device__ __forceinline__ void load4(const float *p, float *v) {
    f32x4 u = *(const f32x4 *)p;
    v[0] = u[0]; v[1] = u[1]; v[2] = u[2]; v[3] = u[3];
}
__device__ __forceinline__ void store4(bf16_t *p, const float *v) {
    u32x2 u = {(unsigned)f2bf(v[0]) | ((unsigned)f2bf(v[1]) << 16), (unsigned)f2bf(v[2]) | ((unsigned)f2bf(v[3]) << 16)};
    *(u32x2 *)p = u;
}
__device__ __forceinline__ void store4(float *p, const float *v) {
    f32x4 u = {v[0], v[1], v[2], v[3]};
    *(f32x4 *)p = u;
}

// full weighting: coarse I = sum over fine 2I + (a, b, c), a, b, c in {-1, 0, 1}, of
// w(a) w(b) w(c) (w(0) = 1/2, w(+-1) = 1/4), zero outside; times scale
template <typename TI, typename TO>
__global__ void __launch_bounds__(256) k_restrict(NetGrid f, NetGrid cg, int C, const TI *__restrict__ in,
                                                  TO *__restrict__ out, float scale) {
    const int C4 = C / 4;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cg.nvox() * C4) return;
    int c4 = (int)(t % C4);
    long v = t / C4;
    int K = (int)(v % cg.nz), J = (int)((v / cg.nz) % cg.ny), I = (int)(v / ((long)cg.nz * cg.ny));
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int a = -1; a <= 1; a++) {
        int i = 2 * I + a;
        if (i < 0 || i >= f.nx) continue;
        float wa = a == 0 ? 0.5f : 0.25f;
        for (int b = -1; b <= 1; b++) {
            int j = 2 * J + b;
            if (j < 0 || j >= f.ny) continue;
            float wb = wa * (b == 0 ? 0.5f : 0.25f);
            for (int c = -1; c <= 1; c++) {
                int k = 2 * K + c;
                if (k < 0 || k >= f.nz) continue;
                float wc = wb * (c == 0 ? 0.5f : 0.25f);
                float x[4];
                load4(in + (((long)i * f.ny + j) * f.nz + k) * C + c4 * 4, x);
                for (int q = 0; q < 4; q++) acc[q] += wc * x[q];
            }
        }
    }
    for (int q = 0; q < 4; q++) acc[q] *= scale;
    store4(out + v * C + c4 * 4, acc);
}

// trilinear prolongation: fine i even -> coarse i/2; odd -> mean of (i-1)/2 and (i+1)/2
// out = bf16(P(coarse) + skip)   or   out = acc + scale P(coarse)
template <typename TI, typename TS, typename TO>
__global__ void __launch_bounds__(256) k_prolong(NetGrid f, NetGrid cg, int C, const TI *__restrict__ cin,
                                                 const TS *add, TO *out, float scale) {
    const int C4 = C / 4;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= f.nvox() * C4) return;
    int c4 = (int)(t % C4);
    long v = t / C4;
    int k = (int)(v % f.nz), j = (int)((v / f.nz) % f.ny), i = (int)(v / ((long)f.nz * f.ny));
    int ia[2] = {i >> 1, (i + 1) >> 1}, ja[2] = {j >> 1, (j + 1) >> 1}, ka[2] = {k >> 1, (k + 1) >> 1};
    int ni = (i & 1) ? 2 : 1, nj = (j & 1) ? 2 : 1, nk = (k & 1) ? 2 : 1;
    float wi = (i & 1) ? 0.5f : 1.0f, wj = (j & 1) ? 0.5f : 1.0f, wk = (k & 1) ? 0.5f : 1.0f;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int a = 0; a < ni; a++)
        for (int b = 0; b < nj; b++)
            for (int c = 0; c < nk; c++) {
                float x[4];
                load4(cin + (((long)ia[a] * cg.ny + ja[b]) * cg.nz + ka[c]) * C + c4 * 4, x);
                float wgt = wi * wj * wk;
                for (int q = 0; q < 4; q++) acc[q] += wgt * x[q];
            }
    float s[4];
    load4(add + v * C + c4 * 4, s);
    for (int q = 0; q < 4; q++) acc[q] = s[q] + scale * acc[q];
    store4(out + v * C + c4 * 4, acc);
}

__global__ void __launch_bounds__(256) k_mask(long n, const float *__restrict__ d, const bf16_t *__restrict__ y,
                                              bf16_t *__restrict__ out) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = f2bf(y[t] != 0 && !(y[t] & 0x8000u) ? d[t] : 0.0f);
}

__global__ void __launch_bounds__(256) k_f32_to_bf16(long n, const float *__restrict__ in, bf16_t *__restrict__ out) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = f2bf(in[t]);
}

__global__ void __launch_bounds__(256) k_bf16_to_f32(long n, const bf16_t *__restrict__ in, float *__restrict__ out) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = bf2f(in[t]);
}

__global__ void __launch_bounds__(256) k_expand_ch0(long nvox, const float *__restrict__ dy, bf16_t *__restrict__ out) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nvox * 16) out[t] = (t % 16) == 0 ? f2bf(dy[t / 16]) : (bf16_t)0;
}

__global__ void __launch_bounds__(256) k_adam(long n, float *__restrict__ w, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, float lr_t, float b1,
                                              float b2, float eps) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    float gr = g[t];
    float mm = b1 * m[t] + (1.f - b1) * gr;
    float vv = b2 * v[t] + (1.f - b2) * gr * gr;
    m[t] = mm;
    v[t] = vv;
    w[t] -= lr_t * mm / (sqrtf(vv) + eps);
}

// ------------------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------------------
static inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

void launch_pack(const float *W, int cin, int cout, int transpose, bf16_t *packed, hipStream_t s) {
    int ci = transpose ? cout : cin, co = transpose ? cin : cout;
    long total = conv_pack_elems(ci, co);
    hipLaunchKernelGGL(k_pack, dim3(nblk(total)), dim3(256), 0, s, W, cin, cout, transpose, ci, co, total, packed);
}

#define CONV_CASE(CI, CO)                                                                                  \
    if (cin == CI && cout == CO) {                                                                         \
        hipLaunchKernelGGL((k_conv<CI, CO>), dim3((unsigned)g.ntiles()), dim3(256), 0, s, g, in, wpack, bias, \
                           out, mode);                                                                     \
        return 0;                                                                                          \
    }

int launch_conv(int cin, int cout, const NetGrid &g, const bf16_t *in, const bf16_t *wpack, const float *bias,
                void *out, int mode, hipStream_t s) {
    CONV_CASE(8, 16)
    CONV_CASE(16, 16)
    CONV_CASE(16, 32)
    CONV_CASE(32, 32)
    CONV_CASE(32, 64)
    CONV_CASE(64, 64)
    CONV_CASE(64, 32)
    CONV_CASE(32, 16)
    return -1;
}

// workgroups per channel pair: 512 workgroups in all fill the 256 CUs at the two workgroups that
// the weight-gradient kernel's LDS allows per CU
int wgrad_slots(int cin, int cout, const NetGrid &g) {
    int npair = (cout / 16) * ((cin + 15) / 16);
    long want = 512 / npair;
    if (want < 8) want = 8;
    return (int)(want < g.ntiles() ? want : g.ntiles());
}

long wgrad_part_elems(int cin, int cout, const NetGrid &g) {
    int npair = (cout / 16) * ((cin + 15) / 16);
    return (long)wgrad_slots(cin, cout, g) * npair * 27 * 256;
}

#define WGRAD_CASE(CI, CO)                                                                                 \
    if (cin == CI && cout == CO) {                                                                         \
        hipLaunchKernelGGL((k_wgrad<CI, CO>), dim3(nslots, npair), dim3(256), 0, s, g, dy, x, part, nslots, bpart); \
        hipLaunchKernelGGL(k_wgrad_final, dim3((unsigned)((27L * cin * cout + 15) / 16)), dim3(256), 0, s, part, cin, \
                           cout, nslots, dW);                                                              \
        if (bpart) hipLaunchKernelGGL(k_bias_final, dim3(1), dim3(256), 0, s, bpart, cout, nslots, db);   \
        return 0;                                                                                          \
    }

int launch_wgrad(int cin, int cout, const NetGrid &g, const bf16_t *dy, const bf16_t *x, float *part, float *dW,
                 float *bpart, float *db, hipStream_t s) {
    if (bpart && !db) return -1;
    int nslots = wgrad_slots(cin, cout, g);
    int npair = (cout / 16) * ((cin + 15) / 16);
    WGRAD_CASE(8, 16)
    WGRAD_CASE(16, 16)
    WGRAD_CASE(16, 32)
    WGRAD_CASE(32, 32)
    WGRAD_CASE(32, 64)
    WGRAD_CASE(64, 64)
    WGRAD_CASE(64, 32)
    WGRAD_CASE(32, 16)
    return -1;
}

void launch_restrict_bf16(const NetGrid &fine, const NetGrid &coarse, int C, const bf16_t *in, bf16_t *out,
                          hipStream_t s) {
    hipLaunchKernelGGL((k_restrict<bf16_t, bf16_t>), dim3(nblk(coarse.nvox() * (C / 4))), dim3(256), 0, s, fine,
                       coarse, C, in, out, 1.0f);
}

void launch_restrict_f32(const NetGrid &fine, const NetGrid &coarse, int C, const float *in, float *out, float scale,
                         hipStream_t s) {
    hipLaunchKernelGGL((k_restrict<float, float>), dim3(nblk(coarse.nvox() * (C / 4))), dim3(256), 0, s, fine, coarse,
                       C, in, out, scale);
}

void launch_prolong_add_bf16(const NetGrid &fine, const NetGrid &coarse, int C, const bf16_t *coarse_in,
                             const bf16_t *skip, bf16_t *out, hipStream_t s) {
    hipLaunchKernelGGL((k_prolong<bf16_t, bf16_t, bf16_t>), dim3(nblk(fine.nvox() * (C / 4))), dim3(256), 0, s, fine,
                       coarse, C, coarse_in, skip, out, 1.0f);
}

void launch_prolong_add_f32(const NetGrid &fine, const NetGrid &coarse, int C, const float *coarse_in,
                            const float *acc, float *out, float scale, hipStream_t s) {
    hipLaunchKernelGGL((k_prolong<float, float, float>), dim3(nblk(fine.nvox() * (C / 4))), dim3(256), 0, s, fine,
                       coarse, C, coarse_in, acc, out, scale);
}

void launch_mask(long n, const float *d, const bf16_t *y, bf16_t *out, hipStream_t s) {
    hipLaunchKernelGGL(k_mask, dim3(nblk(n)), dim3(256), 0, s, n, d, y, out);
}

void launch_f32_to_bf16(long n, const float *in, bf16_t *out, hipStream_t s) {
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(n)), dim3(256), 0, s, n, in, out);
}

void launch_bf16_to_f32(long n, const bf16_t *in, float *out, hipStream_t s) {
    hipLaunchKernelGGL(k_bf16_to_f32, dim3(nblk(n)), dim3(256), 0, s, n, in, out);
}

void launch_expand_ch0(long nvox, const float *dy, bf16_t *out16, hipStream_t s) {
    hipLaunchKernelGGL(k_expand_ch0, dim3(nblk(nvox * 16)), dim3(256), 0, s, nvox, dy, out16);
}

void launch_adam(long n, float *w, const float *g, float *m, float *v, float lr_t, float beta1, float beta2,
                 float eps, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3(nblk(n)), dim3(256), 0, s, n, w, g, m, v, lr_t, beta1, beta2, eps);
}

}  // namespace sr
