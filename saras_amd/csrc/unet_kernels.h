// Kernels of the learned initial guess (a 3D U-Net on node-centred grids, channels-last bf16
// tensors, fp32 accumulation on the matrix cores).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace sr {

typedef uint16_t bf16_t;   // bfloat16 bit pattern

// grid of one network level: node counts and the 4 x 4 x 16 output tiles of the convolutions
struct NetGrid {
    int nx, ny, nz;
    int tx, ty, tz;   // tile counts
    __host__ __device__ long nvox() const { return (long)nx * ny * nz; }
    __host__ __device__ long ntiles() const { return (long)tx * ty * tz; }
};

inline NetGrid net_grid(int nx, int ny, int nz) {
    NetGrid g{nx, ny, nz, (nx + 3) / 4, (ny + 3) / 4, (nz + 15) / 16};
    return g;
}

enum ConvOut { CONV_BF16_RELU = 0, CONV_F32 = 1, CONV_F32_CH0 = 2 };

// packed-weight size (bf16 elements) of a convolution CIN -> COUT
long conv_pack_elems(int cin, int cout);
// W (27, COUT, CIN) fp32 master weights -> packed bf16 fragments of the convolution; with
// transpose, the packing of the adjoint convolution (COUT -> CIN, taps mirrored)
void launch_pack(const float *W, int cin, int cout, int transpose, bf16_t *packed, hipStream_t s);
// out = conv(in) (+ bias, ReLU, bf16) or raw fp32
int launch_conv(int cin, int cout, const NetGrid &g, const bf16_t *in, const bf16_t *wpack, const float *bias,
                void *out, int mode, hipStream_t s);
// weight gradient: partials per slot, then the fixed-order sum into dW (27, COUT, CIN) fp32
int wgrad_slots(int cin, int cout, const NetGrid &g);
long wgrad_part_elems(int cin, int cout, const NetGrid &g);
// with bpart (wgrad_slots x cout floats) the bias gradient db (cout) of the same output gradient
// comes from the same pass: the blocks of input block 0 add the staged gradient tiles per channel
int launch_wgrad(int cin, int cout, const NetGrid &g, const bf16_t *dy, const bf16_t *x, float *part, float *dW,
                 float *bpart, float *db, hipStream_t s);
// transfer operators on channels-last tensors (C a multiple of 4)
void launch_restrict_bf16(const NetGrid &fine, const NetGrid &coarse, int C, const bf16_t *in, bf16_t *out,
                          hipStream_t s);
void launch_restrict_f32(const NetGrid &fine, const NetGrid &coarse, int C, const float *in, float *out, float scale,
                         hipStream_t s);
// out = bf16(P(coarse) + skip)
void launch_prolong_add_bf16(const NetGrid &fine, const NetGrid &coarse, int C, const bf16_t *coarse_in,
                             const bf16_t *skip, bf16_t *out, hipStream_t s);
// out = acc + scale P(coarse)   (fp32; acc may alias out)
void launch_prolong_add_f32(const NetGrid &fine, const NetGrid &coarse, int C, const float *coarse_in,
                            const float *acc, float *out, float scale, hipStream_t s);
// out = bf16(d * (y > 0))
void launch_mask(long n, const float *d, const bf16_t *y, bf16_t *out, hipStream_t s);
void launch_f32_to_bf16(long n, const float *in, bf16_t *out, hipStream_t s);
void launch_bf16_to_f32(long n, const bf16_t *in, float *out, hipStream_t s);
// one channel of dy into a 16-channel bf16 tensor (the other channels zero)
void launch_expand_ch0(long nvox, const float *dy, bf16_t *out16, hipStream_t s);
// Adam on fp32 master weights (bias-corrected step sizes given by the host)
void launch_adam(long n, float *w, const float *g, float *m, float *v, float lr_t, float beta1, float beta2,
                 float eps, hipStream_t s);

}  // namespace sr
