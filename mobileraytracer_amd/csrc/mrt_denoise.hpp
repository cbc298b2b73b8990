// Denoiser kernels (mrt_denoise.hip): BF16 NHWC tensors, 3x3 convolutions on
// MFMA with FP32 accumulation, pooling, upsampling and input / output packing.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mrt {

// One 3x3 convolution (stride 1, zero padding 1) over an H x W frame.
// Input channels [x_off, x_off + cin) of x (x_stride channels per pixel);
// weights w[tap][cout][cin] BF16 (tap = 3 * dy + dx, dy, dx in 0..2 for
// offsets -1..1); FP32 bias (or null).  Output: BF16 channels [y_off, y_off +
// cout) of y (y_stride per pixel), or, when y32 is set, FP32 channels
// [0, y32_channels) of y32 (y32_channels per pixel; the others are dropped).
// cin and cout are multiples of 32; offsets and strides multiples of 8.
struct ConvArgs {
    const uint16_t* x;
    uint32_t x_stride, x_off, cin;
    const uint16_t* w;
    const float* bias;
    uint32_t cout;
    uint16_t* y;
    uint32_t y_stride, y_off;
    float* y32;
    uint32_t y32_channels;
    uint32_t H, W;
    uint32_t relu;
};

hipError_t launch_conv3x3(const ConvArgs& a, hipStream_t stream);
// 2x2 max pooling: channels [x_off, x_off + c) of x (H x W, H and W even) into
// y (H/2 x W/2, c channels per pixel).
hipError_t launch_pool2(const uint16_t* x, uint32_t x_stride, uint32_t x_off, uint32_t c, uint32_t H, uint32_t W,
                        uint16_t* y, hipStream_t stream);
// 2x nearest upsampling: x (H/2 x W/2, c channels) into channels [y_off, y_off
// + c) of y (H x W, y_stride per pixel).
hipError_t launch_up2(const uint16_t* x, uint32_t c, uint32_t H, uint32_t W, uint16_t* y, uint32_t y_stride,
                      uint32_t y_off, hipStream_t stream);
// Network input (32 channels per pixel, 0..9 used) from radiance (3 floats
// per pixel, multiplied by rad_scale) and AOVs (7 floats) of a w x h frame,
// zero-padded to H x W.
hipError_t launch_pack(const float* rad, float rad_scale, const float* aov, uint32_t w, uint32_t h,
                       float depth_scale, uint32_t H, uint32_t W, uint16_t* x, hipStream_t stream);
// Radiance (3 floats per pixel, w x h) from the network output (c FP32
// channels per pixel of a W-wide frame).
hipError_t launch_unpack(const float* y, uint32_t c, uint32_t W, uint32_t w, uint32_t h, float* out,
                         hipStream_t stream);

}  // namespace mrt
