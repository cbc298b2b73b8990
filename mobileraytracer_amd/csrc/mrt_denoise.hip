// Denoiser kernels: BF16 3x3 convolutions on MFMA (FP32 accumulation), 2x2
// max pooling, 2x nearest upsampling and the input / output packing of a
// U-Net over the renderer's radiance and first-hit AOVs.
//
// Tensors are NHWC BF16 (uint16 storage): element (y, x, c) of a buffer with
// `stride` channels per pixel sits at (y * W + x) * stride + c, so a layer can
// read or write a channel slice [off, off + channels) of a wider buffer (the
// skip concatenations are slices of one buffer).
//
// Convolution = implicit GEMM: C[pixel][cout] = sum over the 9 taps and the
// input channels of X[neighbour pixel][cin] * Wt[tap][cout][cin].  One wave
// computes 64 consecutive pixels of a row by 32 output channels as 4 x 2
// tiles of v_mfma_f32_16x16x32_bf16 (16 pixels x 16 channels, K = 32 input
// channels of one tap): lane l holds pixel / output channel (l % 16) and the 8
// consecutive input channels 8 * (l / 16) .. + 7 of the K step, one 16-byte
// load each; the 4 accumulators of a tile are output rows 4 * (l / 16) .. + 3
// of column l % 16.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mrt_denoise.hpp"

namespace mrt {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// Round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t f_to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    return static_cast<uint16_t>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) {
    bf16x8 r;
    __builtin_memcpy(&r, &v, sizeof(r));
    return r;
}

__global__ __launch_bounds__(64) void k_conv3x3(ConvArgs a) {
    const uint32_t lane = threadIdx.x;
    const uint32_t x0 = blockIdx.x * 64u;
    const uint32_t y = blockIdx.y;
    const uint32_t n0 = blockIdx.z * 32u;
    const uint32_t li = lane & 15u, lk = (lane >> 4) * 8u;
    floatx4 acc[4][2];
    for (int m = 0; m < 4; ++m)
        for (int n = 0; n < 2; ++n) acc[m][n] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    for (int tap = 0; tap < 9; ++tap) {
        const int yy = static_cast<int>(y) + tap / 3 - 1;
        const bool rowIn = yy >= 0 && yy < static_cast<int>(a.H);
        const uint16_t* wt = a.w + (static_cast<size_t>(tap) * a.cout + n0 + li) * a.cin + lk;
        for (uint32_t kc = 0; kc < a.cin; kc += 32u) {
            uint4 av[4];
            for (int m = 0; m < 4; ++m) {
                const int xx = static_cast<int>(x0 + 16u * m + li) + tap % 3 - 1;
                const bool in = rowIn && xx >= 0 && xx < static_cast<int>(a.W);
                av[m] = make_uint4(0u, 0u, 0u, 0u);
                if (in)
                    av[m] = *reinterpret_cast<const uint4*>(
                        a.x + (static_cast<size_t>(yy) * a.W + xx) * a.x_stride + a.x_off + kc + lk);
            }
            const uint4 b0 = *reinterpret_cast<const uint4*>(wt + kc);
            const uint4 b1 = *reinterpret_cast<const uint4*>(wt + static_cast<size_t>(16u) * a.cin + kc);
            for (int m = 0; m < 4; ++m) {
                acc[m][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(av[m]), as_bf16x8(b0), acc[m][0], 0, 0, 0);
                acc[m][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(av[m]), as_bf16x8(b1), acc[m][1], 0, 0, 0);
            }
        }
    }
    for (int n = 0; n < 2; ++n) {
        const uint32_t co = n0 + 16u * n + li;
        const float b = a.bias ? a.bias[co] : 0.0f;
        for (int m = 0; m < 4; ++m) {
            for (int v = 0; v < 4; ++v) {
                const uint32_t x = x0 + 16u * m + (lane >> 4) * 4u + v;
                if (x >= a.W) continue;
                float r = acc[m][n][v] + b;
                if (a.relu) r = fmaxf(r, 0.0f);
                const size_t px = static_cast<size_t>(y) * a.W + x;
                if (a.y32) {
                    if (co < a.y32_channels) a.y32[px * a.y32_channels + co] = r;
                } else {
                    a.y[px * a.y_stride + a.y_off + co] = f_to_bf16(r);
                }
            }
        }
    }
}

// 2x2 max pooling of channels [off, off + c) of x (H x W) into y (H/2 x W/2,
// c channels).
__global__ void k_pool2(const uint16_t* __restrict__ x, uint32_t x_stride, uint32_t x_off, uint32_t c, uint32_t H,
                        uint32_t W, uint16_t* __restrict__ y) {
    const uint32_t W2 = W / 2u, H2 = H / 2u;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<size_t>(H2) * W2 * c) return;
    const uint32_t ch = i % c;
    const size_t p = i / c;
    const uint32_t px = p % W2, py = static_cast<uint32_t>(p / W2);
    float m = -INFINITY;
    for (uint32_t dy = 0; dy < 2; ++dy)
        for (uint32_t dx = 0; dx < 2; ++dx)
            m = fmaxf(m, bf16_to_f(x[((2u * py + dy) * static_cast<size_t>(W) + 2u * px + dx) * x_stride + x_off + ch]));
    y[i] = f_to_bf16(m);
}

// 2x nearest upsampling of x (H/2 x W/2, c channels) into channels [off, off +
// c) of y (H x W).
__global__ void k_up2(const uint16_t* __restrict__ x, uint32_t c, uint32_t H, uint32_t W, uint16_t* __restrict__ y,
                      uint32_t y_stride, uint32_t y_off) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<size_t>(H) * W * c) return;
    const uint32_t ch = i % c;
    const size_t p = i / c;
    const uint32_t px = p % W, py = static_cast<uint32_t>(p / W);
    y[p * y_stride + y_off + ch] = x[((py / 2u) * static_cast<size_t>(W / 2u) + px / 2u) * c + ch];
}

// Network input: radiance * rad_scale (3 floats) as log(1 + max(x, 0)),
// albedo, normal and depth * depth_scale (7 floats) as BF16 channels 0..9 of
// a 32-channel NHWC tensor (channels 10..31 zero).  The frame (w x h) sits in
// the top-left corner of the zero-padded H x W tensor.
__global__ void k_pack(const float* __restrict__ rad, float rad_scale, const float* __restrict__ aov, uint32_t w,
                       uint32_t h, float depth_scale, uint32_t H, uint32_t W, uint16_t* __restrict__ x) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<size_t>(H) * W) return;
    const uint32_t px = i % W, py = static_cast<uint32_t>(i / W);
    uint16_t* o = x + i * 32u;
    float v[10] = {};
    if (px < w && py < h) {
        const size_t s = static_cast<size_t>(py) * w + px;
        for (int k = 0; k < 3; ++k) v[k] = log1pf(fmaxf(rad[3 * s + k] * rad_scale, 0.0f));
        for (int k = 0; k < 6; ++k) v[3 + k] = aov[7 * s + k];
        v[9] = aov[7 * s + 6] * depth_scale;
    }
    for (int k = 0; k < 32; ++k) o[k] = k < 10 ? f_to_bf16(v[k]) : static_cast<uint16_t>(0);
}

// Network output (FP32, `c` channels per pixel, channels 0..2 used) back to
// radiance: exp(y) - 1, clamped at 0, cropped to w x h.
__global__ void k_unpack(const float* __restrict__ y, uint32_t c, uint32_t W, uint32_t w, uint32_t h,
                         float* __restrict__ out) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<size_t>(w) * h) return;
    const uint32_t px = i % w, py = static_cast<uint32_t>(i / w);
    const float* s = y + (static_cast<size_t>(py) * W + px) * c;
    for (int k = 0; k < 3; ++k) out[3 * i + k] = fmaxf(expm1f(s[k]), 0.0f);
}

uint32_t blocks_for(size_t n, uint32_t b) { return static_cast<uint32_t>((n + b - 1) / b); }

}  // namespace

hipError_t launch_conv3x3(const ConvArgs& a, hipStream_t stream) {
    if (a.cin % 32u || a.cout % 32u || a.x_off % 8u || a.x_stride % 8u || a.x_off + a.cin > a.x_stride)
        return hipErrorInvalidValue;
    if (!a.y32 && (a.y_off % 8u || a.y_stride % 8u || a.y_off + a.cout > a.y_stride)) return hipErrorInvalidValue;
    dim3 grid((a.W + 63u) / 64u, a.H, a.cout / 32u);
    k_conv3x3<<<grid, 64, 0, stream>>>(a);
    return hipGetLastError();
}

hipError_t launch_pool2(const uint16_t* x, uint32_t x_stride, uint32_t x_off, uint32_t c, uint32_t H, uint32_t W,
                        uint16_t* y, hipStream_t stream) {
    const size_t n = static_cast<size_t>(H / 2u) * (W / 2u) * c;
    k_pool2<<<blocks_for(n, 256), 256, 0, stream>>>(x, x_stride, x_off, c, H, W, y);
    return hipGetLastError();
}

hipError_t launch_up2(const uint16_t* x, uint32_t c, uint32_t H, uint32_t W, uint16_t* y, uint32_t y_stride,
                      uint32_t y_off, hipStream_t stream) {
    const size_t n = static_cast<size_t>(H) * W * c;
    k_up2<<<blocks_for(n, 256), 256, 0, stream>>>(x, c, H, W, y, y_stride, y_off);
    return hipGetLastError();
}

hipError_t launch_pack(const float* rad, float rad_scale, const float* aov, uint32_t w, uint32_t h,
                       float depth_scale, uint32_t H, uint32_t W, uint16_t* x, hipStream_t stream) {
    k_pack<<<blocks_for(static_cast<size_t>(H) * W, 256), 256, 0, stream>>>(rad, rad_scale, aov, w, h, depth_scale, H,
                                                                           W, x);
    return hipGetLastError();
}

hipError_t launch_unpack(const float* y, uint32_t c, uint32_t W, uint32_t w, uint32_t h, float* out,
                         hipStream_t stream) {
    k_unpack<<<blocks_for(static_cast<size_t>(w) * h, 256), 256, 0, stream>>>(y, c, W, w, h, out);
    return hipGetLastError();
}

}  // namespace mrt
