// Denoiser kernels: BF16 3x3 convolutions on MFMA (FP32 accumulation) and
// 2x2 max pooling for a U-Net over the renderer's radiance and first-hit
// AOVs.
//
// Tensors are NHWC BF16 (uint16 storage): element (y, x, c) of a buffer with
// `stride` channels per pixel sits at (y * W + x) * stride + c.
//
// Convolution = implicit GEMM: C[cout][pixel] = sum over the 9 taps and the
// input channels of Wt[tap][cout][cin] * X[neighbour pixel][cin], on
// v_mfma_f32_16x16x32_bf16 (16 output channels x 16 pixels, K = 32 input
// channels of one tap).  A 256-thread workgroup computes a tile of 4 rows x
// 64 pixels x NT output channels (wave w: row w, NT/16 x 4 MFMA tiles).  Per
// 32-channel chunk the input rows of the tile with a one-pixel halo (6 x 66
// pixels) are staged in LDS once for the 4 waves, 80 bytes per pixel (64
// bytes of channels + 16 of padding: consecutive pixels then start 20 banks
// apart, so the 16-byte fragment reads of 8 or 16 lanes hit distinct banks).
// The chunk's weights (9 taps x NT output channels) go to LDS the same way
// for NT = 64; for NT = 32 each wave loads them into registers (72 VGPRs),
// which leaves the LDS to the input tile and more workgroups per CU.  Lane l
// holds output channel / pixel (l % 16) and input channels 8 * (l / 16) .. + 7
// of a K step; the 4 accumulators of a tile are output channels
// 4 * (l / 16) .. + 3 of pixel l % 16, stored as one 8-byte write.
//
// The decoder's concatenation [upsampled coarser level, skip] is never
// written: chunks of the upsampled part are staged from the coarser tensor at
// (y / 2, x / 2).  The network input is packed from FP32 radiance and AOVs
// while staging the first convolution, and the last convolution writes FP32
// radiance.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mrt_denoise.hpp"

namespace mrt {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

__device__ __forceinline__ float bf16_lo(uint32_t v) { return __uint_as_float(v << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t v) { return __uint_as_float(v & 0xFFFF0000u); }

// Round to nearest even (finite inputs).
__device__ __forceinline__ uint32_t f_to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}

__device__ __forceinline__ uint32_t pack2(float lo, float hi) { return f_to_bf16(lo) | (f_to_bf16(hi) << 16); }

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) {
    bf16x8 r;
    __builtin_memcpy(&r, &v, sizeof(r));
    return r;
}

constexpr uint32_t kTileRows = 4, kTileW = 64, kHaloW = kTileW + 2, kRowQuads = 5;
constexpr uint32_t kTileQuads = (kTileRows + 2) * kHaloW * kRowQuads;

// Stages the packed network input (32 channels, 0..9 used) of the tile rows
// and halo: log(1 + max(rad * rad_scale, 0)), aov[0..5], aov[6] *
// depth_scale; zero outside the fw x fh frame.  The loads of both rounds of
// pixels are issued before any value is converted.
__device__ __forceinline__ void stage_packed(const ConvArgs& a, uint4* tile, uint32_t tid, int x0, int y0) {
    constexpr uint32_t kPx = (kTileRows + 2) * kHaloW;
    constexpr uint32_t kIters = (kPx + 255u) / 256u;
    float v[kIters][10];
    bool inside[kIters];
#pragma unroll
    for (uint32_t it = 0; it < kIters; ++it) {
        const uint32_t px = tid + it * 256u;
        const uint32_t r = px / kHaloW, c = px % kHaloW;
        const int yy = y0 - 1 + static_cast<int>(r), xx = x0 - 1 + static_cast<int>(c);
        const bool in = px < kPx && yy >= 0 && xx >= 0 && yy < static_cast<int>(a.fh) && xx < static_cast<int>(a.fw);
        inside[it] = in;
        const size_t s = in ? static_cast<size_t>(yy) * a.fw + xx : 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) v[it][k] = in ? a.rad[3 * s + k] : 0.0f;
#pragma unroll
        for (int k = 0; k < 7; ++k) v[it][3 + k] = in ? a.aov[7 * s + k] : 0.0f;
    }
#pragma unroll
    for (uint32_t it = 0; it < kIters; ++it) {
        const uint32_t px = tid + it * 256u;
        if (px >= kPx) continue;
        float w[10] = {};
        if (inside[it]) {
#pragma unroll
            for (int k = 0; k < 3; ++k) w[k] = log1pf(fmaxf(v[it][k] * a.rad_scale, 0.0f));
#pragma unroll
            for (int k = 0; k < 6; ++k) w[3 + k] = v[it][3 + k];
            w[9] = v[it][9] * a.depth_scale;
        }
        uint4* o = tile + px * kRowQuads;
        o[0] = make_uint4(pack2(w[0], w[1]), pack2(w[2], w[3]), pack2(w[4], w[5]), pack2(w[6], w[7]));
        o[1] = make_uint4(pack2(w[8], w[9]), 0u, 0u, 0u);
        o[2] = make_uint4(0u, 0u, 0u, 0u);
        o[3] = make_uint4(0u, 0u, 0u, 0u);
    }
}

template <int NT>
__global__ __launch_bounds__(256) void k_conv3x3(ConvArgs a) {
    constexpr int NJ = NT / 16;
    constexpr bool kWeightsLds = NT > 32;
    __shared__ uint4 smem[kTileQuads + (kWeightsLds ? 9 * NT * kRowQuads : 0)];
    uint4* tile = smem;
    uint4* wts = smem + kTileQuads;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6;
    const uint32_t li = lane & 15u, lg = lane >> 4;
    const int x0 = static_cast<int>(blockIdx.x * kTileW), y0 = static_cast<int>(blockIdx.y * kTileRows);
    const uint32_t n0 = blockIdx.z * NT;
    const int H = static_cast<int>(a.H), W = static_cast<int>(a.W);
    floatx4 acc[NJ][4];
#pragma unroll
    for (int j = 0; j < NJ; ++j)
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[j][m] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    // Staging: all of a chunk's loads are issued before any of them is
    // written to LDS, and the next chunk's loads are in flight while the
    // current chunk's MFMA loop runs.
    constexpr uint32_t kStageQuads = (kTileRows + 2) * kHaloW * 4u;
    constexpr uint32_t kStageIters = (kStageQuads + 255u) / 256u;
    constexpr uint32_t kWQuads = 9u * NT * 4u;
    constexpr uint32_t kWIters = kWeightsLds ? (kWQuads + 255u) / 256u : 1u;
    uint4 sv[kStageIters];
    uint4 wq[kWIters];
    auto load_input = [&](uint32_t kc) {
        const bool up = kc < a.cu;
#pragma unroll
        for (uint32_t it = 0; it < kStageIters; ++it) {
            const uint32_t p = tid + it * 256u;
            sv[it] = make_uint4(0u, 0u, 0u, 0u);
            if (p >= kStageQuads) continue;
            const uint32_t q = p & 3u, px = p >> 2;
            const uint32_t r = px / kHaloW, c = px % kHaloW;
            const int yy = y0 - 1 + static_cast<int>(r), xx = x0 - 1 + static_cast<int>(c);
            if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                const uint16_t* src =
                    up ? a.xu + (static_cast<size_t>(yy >> 1) * (a.W >> 1) + (xx >> 1)) * a.cu + kc
                       : a.x + (static_cast<size_t>(yy) * a.W + xx) * a.x_stride + a.x_off + (kc - a.cu);
                sv[it] = *reinterpret_cast<const uint4*>(src + q * 8u);
            }
        }
    };
    auto load_weights_lds = [&](uint32_t kc) {
#pragma unroll
        for (uint32_t it = 0; it < kWIters; ++it) {
            const uint32_t p = tid + it * 256u;
            wq[it] = make_uint4(0u, 0u, 0u, 0u);
            if (p >= kWQuads) continue;
            const uint32_t q = p & 3u, row = p >> 2;  // row = tap * NT + n
            const uint32_t tap = row / NT, n = row % NT;
            wq[it] = *reinterpret_cast<const uint4*>(a.w + (static_cast<size_t>(tap) * a.cout + n0 + n) * a.cin + kc +
                                                     q * 8u);
        }
    };
    if (!a.rad) load_input(0);
    if constexpr (kWeightsLds) load_weights_lds(0);
    for (uint32_t kc = 0; kc < a.cin; kc += 32u) {
        uint4 wr[kWeightsLds ? 1 : 9][NJ];
        if constexpr (!kWeightsLds) {
            const uint16_t* wk = a.w + static_cast<size_t>(n0 + li) * a.cin + kc + lg * 8u;
#pragma unroll
            for (int tap = 0; tap < 9; ++tap)
#pragma unroll
                for (int j = 0; j < NJ; ++j)
                    wr[tap][j] =
                        *reinterpret_cast<const uint4*>(wk + (static_cast<size_t>(tap) * a.cout + 16u * j) * a.cin);
        }
        __syncthreads();
        if (a.rad) {
            stage_packed(a, tile, tid, x0, y0);
        } else {
#pragma unroll
            for (uint32_t it = 0; it < kStageIters; ++it) {
                const uint32_t p = tid + it * 256u;
                if (p < kStageQuads) tile[(p >> 2) * kRowQuads + (p & 3u)] = sv[it];
            }
        }
        if constexpr (kWeightsLds) {
#pragma unroll
            for (uint32_t it = 0; it < kWIters; ++it) {
                const uint32_t p = tid + it * 256u;
                if (p < kWQuads) wts[(p >> 2) * kRowQuads + (p & 3u)] = wq[it];
            }
        }
        __syncthreads();
        if (kc + 32u < a.cin) {
            if (!a.rad) load_input(kc + 32u);
            if constexpr (kWeightsLds) load_weights_lds(kc + 32u);
        }
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;
            uint4 bv[NJ];
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                if constexpr (kWeightsLds)
                    bv[j] = wts[(tap * NT + 16 * j + li) * kRowQuads + lg];
                else
                    bv[j] = wr[tap][j];
            }
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const uint4 av = tile[((wv + dy) * kHaloW + 16u * m + li + dx) * kRowQuads + lg];
#pragma unroll
                for (int j = 0; j < NJ; ++j)
                    acc[j][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(bv[j]), as_bf16x8(av), acc[j][m], 0,
                                                                        0, 0);
            }
        }
    }
    const int y = y0 + static_cast<int>(wv);
    if (y >= H) return;
    if (a.rad_out) {
        // Output channels 0..2 (lanes 0..15 of the first channel tile):
        // radiance.
        if (n0 != 0 || lg != 0u || y >= static_cast<int>(a.fh)) return;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int x = x0 + 16 * m + static_cast<int>(li);
            if (x >= static_cast<int>(a.fw)) continue;
            float* o = a.rad_out + (static_cast<size_t>(y) * a.fw + x) * 3;
#pragma unroll
            for (int v = 0; v < 3; ++v) o[v] = fmaxf(expm1f(acc[0][m][v] + (a.bias ? a.bias[v] : 0.0f)), 0.0f);
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        const uint32_t co = n0 + 16u * j + 4u * lg;
        float b[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) b[v] = a.bias ? a.bias[co + v] : 0.0f;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const int x = x0 + 16 * m + static_cast<int>(li);
            if (x >= W) continue;
            float r[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                r[v] = acc[j][m][v] + b[v];
                if (a.relu) r[v] = fmaxf(r[v], 0.0f);
            }
            *reinterpret_cast<uint2*>(a.y + (static_cast<size_t>(y) * a.W + x) * a.y_stride + a.y_off + co) =
                make_uint2(pack2(r[0], r[1]), pack2(r[2], r[3]));
        }
    }
}

// 2x2 max pooling, 8 channels (16 bytes) per thread.
__global__ void k_pool2(const uint16_t* __restrict__ x, uint32_t c, uint32_t H, uint32_t W, uint16_t* __restrict__ y) {
    const uint32_t W2 = W / 2u, H2 = H / 2u, cq = c / 8u;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<size_t>(H2) * W2 * cq) return;
    const uint32_t q = i % cq;
    const size_t p = i / cq;
    const uint32_t px = p % W2, py = static_cast<uint32_t>(p / W2);
    uint4 s[4];
    for (uint32_t k = 0; k < 4; ++k)
        s[k] = *reinterpret_cast<const uint4*>(
            x + ((2u * py + (k >> 1)) * static_cast<size_t>(W) + 2u * px + (k & 1u)) * c + q * 8u);
    uint32_t o[4];
    for (int e = 0; e < 4; ++e) {
        const uint32_t v0 = (&s[0].x)[e], v1 = (&s[1].x)[e], v2 = (&s[2].x)[e], v3 = (&s[3].x)[e];
        const float lo = fmaxf(fmaxf(bf16_lo(v0), bf16_lo(v1)), fmaxf(bf16_lo(v2), bf16_lo(v3)));
        const float hi = fmaxf(fmaxf(bf16_hi(v0), bf16_hi(v1)), fmaxf(bf16_hi(v2), bf16_hi(v3)));
        // Maxima of BF16 values are BF16 values: no rounding.
        o[e] = (__float_as_uint(lo) >> 16) | (__float_as_uint(hi) & 0xFFFF0000u);
    }
    *reinterpret_cast<uint4*>(y + i * 8u) = make_uint4(o[0], o[1], o[2], o[3]);
}

template <int NT>
void conv_launch(const ConvArgs& a, dim3 grid, hipStream_t stream) {
    k_conv3x3<NT><<<grid, 256, 0, stream>>>(a);
}

}  // namespace

hipError_t launch_conv3x3(const ConvArgs& a, uint32_t min_blocks, hipStream_t stream) {
    if (a.cin == 0 || a.cin % 32u || a.cout % 32u || a.cu % 32u || a.cu > a.cin || a.x_off % 8u || a.x_stride % 8u)
        return hipErrorInvalidValue;
    if (a.rad ? (a.cin != 32u || a.cu != 0u || !a.aov || a.fw > a.W || a.fh > a.H)
              : (a.cin > a.cu && (!a.x || a.x_off + (a.cin - a.cu) > a.x_stride)))
        return hipErrorInvalidValue;
    if (a.cu && (a.H % 2u || a.W % 2u || !a.xu)) return hipErrorInvalidValue;
    if (a.rad_out ? (a.fw > a.W || a.fh > a.H)
                  : (!a.y || a.y_off % 4u || a.y_stride % 4u || a.y_off + a.cout > a.y_stride))
        return hipErrorInvalidValue;
    const uint32_t tiles = ((a.W + kTileW - 1) / kTileW) * ((a.H + kTileRows - 1) / kTileRows);
    const uint32_t nt = (a.cout % 64u == 0 && tiles * (a.cout / 64u) >= min_blocks) ? 64u : 32u;
    dim3 grid((a.W + kTileW - 1) / kTileW, (a.H + kTileRows - 1) / kTileRows, a.cout / nt);
    if (nt == 64u)
        conv_launch<64>(a, grid, stream);
    else
        conv_launch<32>(a, grid, stream);
    return hipGetLastError();
}

hipError_t launch_pool2(const uint16_t* x, uint32_t c, uint32_t H, uint32_t W, uint16_t* y, hipStream_t stream) {
    if (c % 8u || H % 2u || W % 2u) return hipErrorInvalidValue;
    const size_t n = static_cast<size_t>(H / 2u) * (W / 2u) * (c / 8u);
    k_pool2<<<static_cast<uint32_t>((n + 255) / 256), 256, 0, stream>>>(x, c, H, W, y);
    return hipGetLastError();
}

}  // namespace mrt
