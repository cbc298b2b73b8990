// Denoiser kernels (mrt_denoise.hip): BF16 NHWC tensors, 3x3 convolutions on
// MFMA with FP32 accumulation and 2x2 max pooling.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mrt {

// One 3x3 convolution (stride 1, zero padding 1) over an H x W frame.
//
// Input, cin channels (a multiple of 32), in 32-channel chunks:
//   * channels [0, cu) (cu a multiple of 32, possibly 0): xu, an H/2 x W/2
//     tensor of cu channels per pixel, upsampled 2x (nearest);
//   * channels [cu, cin): channels [x_off, x_off + cin - cu) of x (x_stride
//     channels per pixel);
//   * or, when rad is set (cin = 32, cu = 0): the network input packed from a
//     frame of fw x fh pixels at the top left of the H x W frame, zero
//     elsewhere: log(1 + max(rad * rad_scale, 0)) (3 channels), aov[0..5],
//     aov[6] * depth_scale, then 22 zero channels.
// Weights w[tap][cout][cin] BF16 (tap = 3 * dy + dx for offsets dy - 1,
// dx - 1); FP32 bias (or null); ReLU when relu is set.
//
// Output (cout a multiple of 32):
//   * BF16 channels [y_off, y_off + cout) of y (y_stride per pixel);
//   * or, when rad_out is set: FP32 radiance max(exp(v) - 1, 0) of output
//     channels 0..2 for the pixels of the fw x fh frame (3 floats per pixel,
//     row length fw); the other channels are dropped.
// Offsets and strides are multiples of 8.
struct ConvArgs {
    const uint16_t* x;
    uint32_t x_stride, x_off, cin;
    const uint16_t* xu;
    uint32_t cu;
    const float* rad;
    const float* aov;
    float rad_scale, depth_scale;
    const uint16_t* w;
    const float* bias;
    uint32_t cout;
    uint16_t* y;
    uint32_t y_stride, y_off;
    float* rad_out;
    uint32_t fw, fh;
    uint32_t H, W;
    uint32_t relu;
};

// min_blocks: the channel block per workgroup (32 or 64 output channels) is
// the widest that still gives this many workgroups.
hipError_t launch_conv3x3(const ConvArgs& a, uint32_t min_blocks, hipStream_t stream);
// 2x2 max pooling: x (H x W, c channels per pixel, H and W even, c a
// multiple of 8) into y (H/2 x W/2, c channels per pixel).
hipError_t launch_pool2(const uint16_t* x, uint32_t c, uint32_t H, uint32_t W, uint16_t* y, hipStream_t stream);

}  // namespace mrt
