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

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

// Two floats to packed BF16 (round to nearest even: v_cvt_pk_bf16_f32), lo in
// the low half.
__device__ __forceinline__ uint32_t pack2(float lo, float hi) {
    const bf16x2 b = __builtin_convertvector(floatx2{lo, hi}, bf16x2);
    uint32_t u;
    __builtin_memcpy(&u, &b, sizeof(u));
    return u;
}

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) {
    bf16x8 r;
    __builtin_memcpy(&r, &v, sizeof(r));
    return r;
}

// 16-byte load at a byte offset from a (wave-uniform) base.
__device__ __forceinline__ uint4 ld16(const uint16_t* base, uint32_t byte_off) {
    return *reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(base) + byte_off);
}

// Wave-uniform value in a scalar register.
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

constexpr uint32_t kTileRows = 4, kTileW = 64, kHaloW = kTileW + 2, kRowQuads = 5;
constexpr uint32_t kTileQuads = (kTileRows + 2) * kHaloW * kRowQuads;

// What a convolution reads and writes: BF16 tensors; the packed network input
// (first layer); FP32 radiance out (last layer); BF16 tensors with the output
// masked by a ReLU's output (training's backward-data convolutions).
enum : int { kModeTensor = 0, kModePackedIn = 1, kModeRadOut = 2, kModeMasked = 3 };

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

// Addressing: every row of the tile (input row, output row of a wave) is a
// wave-uniform base pointer in scalar registers plus a 32-bit per-lane
// element offset within the row, so that loads and stores take the
// scalar-base + vector-offset form and no lane does 64-bit index arithmetic.
// Occupancy: NT = 32 workgroups (31.7 KB of LDS) are held to 168 registers
// per lane, 3 waves per SIMD (2 with the compiler's default register use);
// NT = 64 workgroups (77.8 KB of LDS) fit twice per compute unit anyway.
template <int NT, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NT == 32 ? 3 : 2))) void k_conv3x3(ConvArgs a) {
    constexpr int NJ = NT / 16;
    constexpr bool kWeightsLds = NT > 32;
    static_assert(NT == 32 || NT == 64, "32 or 64 output channels per workgroup");
    __shared__ uint4 smem[kTileQuads + (kWeightsLds ? 9 * NT * kRowQuads : 0)];
    uint4* tile = smem;
    uint4* wts = smem + kTileQuads;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = uniform(tid >> 6);
    const uint32_t li = lane & 15u, lg = lane >> 4;
    const int x0 = static_cast<int>(blockIdx.x * kTileW), y0 = static_cast<int>(blockIdx.y * kTileRows);
    // blockIdx.z = frame * (cout / NT) + channel block.
    const uint32_t nblk = uniform(a.cout / NT);
    const uint32_t img = uniform(blockIdx.z / nblk);
    const uint32_t n0 = (blockIdx.z - img * nblk) * NT;
    const int H = static_cast<int>(a.H), W = static_cast<int>(a.W);
    const size_t plane = static_cast<size_t>(a.H) * a.W;
    const uint16_t* x_img = a.x + img * plane * a.x_stride;
    const uint16_t* xu_img = a.xu + img * (plane / 4u) * a.cu;
    // Accumulators start at the bias: output channels n0 + 16 j + 4 lg + v.
    floatx4 acc[NJ][4];
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
        floatx4 b{0.0f, 0.0f, 0.0f, 0.0f};
        if (a.bias) {
            const float* bj = a.bias + n0 + 16u * j + 4u * lg;
            b = floatx4{bj[0], bj[1], bj[2], bj[3]};
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[j][m] = b;
    }
    // Input staging: thread t carries quad (t & 3) of halo column t >> 2 of
    // each of the 6 tile rows; threads 0..47 also carry the 2 remaining
    // columns (quad t & 3 of column 64 + ((t >> 2) & 1) of row t >> 3).  All
    // of a chunk's loads are issued before any of them is written to LDS, and
    // the next chunk's loads are in flight while the current chunk's MFMA
    // loop runs.
    constexpr uint32_t kRows = kTileRows + 2;
    const uint32_t q = tid & 3u, cm = tid >> 2;
    const int xm = x0 - 1 + static_cast<int>(cm);
    const bool xm_in = xm >= 0 && xm < W;
    const bool has_e = tid < 8u * kRows;
    const uint32_t re = tid >> 3, ce = kTileW + ((tid >> 2) & 1u);
    const int xe = x0 - 1 + static_cast<int>(ce), ye = y0 - 1 + static_cast<int>(re);
    const bool e_in = has_e && xe < W && ye >= 0 && ye < H;
    // Per-lane element offsets within a row (tensor part, upsampled part).
    const uint32_t xs = a.x_stride, cu = a.cu;
    const uint32_t om = 2u * ((xm_in ? static_cast<uint32_t>(xm) : 0u) * xs + 8u * q);
    const uint32_t oum = 2u * ((xm_in ? static_cast<uint32_t>(xm) >> 1 : 0u) * cu + 8u * q);
    const uint32_t oe = 2u * ((e_in ? static_cast<uint32_t>(xe) : 0u) * xs + 8u * q);
    const uint32_t oue = 2u * ((e_in ? static_cast<uint32_t>(xe) >> 1 : 0u) * cu + 8u * q);
    const uint32_t ye_c = e_in ? static_cast<uint32_t>(ye) : 0u;
    uint4 sv[kRows + 1];
    auto load_input = [&](uint32_t kc) {
        const bool up = kc < cu;
        const uint16_t* xb = up ? xu_img + kc : x_img + a.x_off + (kc - cu);
        const size_t row_el = up ? static_cast<size_t>(a.W >> 1) * cu : static_cast<size_t>(a.W) * xs;
        const uint32_t o_m = up ? oum : om;
#pragma unroll
        for (uint32_t r = 0; r < kRows; ++r) {
            const int yy = y0 - 1 + static_cast<int>(r);
            sv[r] = make_uint4(0u, 0u, 0u, 0u);
            if (yy >= 0 && yy < H) {
                const uint16_t* rb = xb + static_cast<size_t>(up ? yy >> 1 : yy) * row_el;
                if (xm_in) sv[r] = ld16(rb, o_m);
            }
        }
        sv[kRows] = make_uint4(0u, 0u, 0u, 0u);
        if (e_in)
            sv[kRows] = ld16(xb + static_cast<size_t>(up ? ye_c >> 1 : ye_c) * row_el, up ? oue : oe);
    };
    // Weights of a chunk, each 16-channel block of a tap one contiguous 1 KB
    // of the fragment-order copy.  NT = 64: LDS, wave w staging block w of
    // every tap (lane (li, lg) into row tap * 64 + 16 w + li, quad lg);
    // NT = 32: registers, lane (li, lg) holding cin 8 lg .. + 7 of output
    // channel n0 + 16 j + li per tap.
    const uint32_t cin = a.cin, cout = a.cout;
    const uint16_t* wfb = a.wf + static_cast<size_t>(n0 >> 4) * (cin >> 5) * 512u;
    const size_t wf_tap = static_cast<size_t>(cout >> 4) * (cin >> 5) * 512u, wf_blk = (cin >> 5) * 512u;
    constexpr uint32_t kWIters = kWeightsLds ? 9u : 1u;
    uint4 wq[kWIters];
    auto load_weights_lds = [&](uint32_t kc) {
#pragma unroll
        for (uint32_t tap = 0; tap < kWIters; ++tap)
            wq[tap] = ld16(wfb + tap * wf_tap + wv * wf_blk + (kc >> 5) * 512u, 16u * lane);
    };
    if constexpr (MODE != kModePackedIn) load_input(0);
    if constexpr (kWeightsLds) load_weights_lds(0);
    for (uint32_t kc = 0; kc < cin; kc += 32u) {
        uint4 wr[kWeightsLds ? 1 : 9][NJ];
        if constexpr (!kWeightsLds) {
#pragma unroll
            for (int tap = 0; tap < 9; ++tap)
#pragma unroll
                for (int j = 0; j < NJ; ++j)
                    wr[tap][j] = ld16(wfb + tap * wf_tap + j * wf_blk + (kc >> 5) * 512u, 16u * lane);
        }
        __syncthreads();
        if constexpr (MODE == kModePackedIn) {
            stage_packed(a, tile, tid, x0, y0);
        } else {
#pragma unroll
            for (uint32_t r = 0; r < kRows; ++r) tile[(r * kHaloW + cm) * kRowQuads + q] = sv[r];
            if (has_e) tile[(re * kHaloW + ce) * kRowQuads + q] = sv[kRows];
        }
        if constexpr (kWeightsLds) {
#pragma unroll
            for (uint32_t tap = 0; tap < kWIters; ++tap) wts[(tap * NT + 16u * wv + li) * kRowQuads + lg] = wq[tap];
        }
        __syncthreads();
        if (kc + 32u < cin) {
            if constexpr (MODE != kModePackedIn) load_input(kc + 32u);
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
    if constexpr (MODE == kModeTensor) {
        if (a.pool) {
            // The 2x2 max pooling of this (ReLU) output into a.pool, as
            // k_pool2 would compute it from the stored tensor (the maximum
            // commutes with the BF16 rounding): pixel pairs from lanes li and
            // li ^ 1, row pairs from waves 2k + 1 through the LDS tile buffer.
            static_assert(NJ * 4 * 128 * sizeof(floatx4) <= sizeof(smem), "pooling exchange fits the LDS");
            floatx4 hp[NJ][4];
#pragma unroll
            for (int j = 0; j < NJ; ++j)
#pragma unroll
                for (int m = 0; m < 4; ++m)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const float r = a.relu ? fmaxf(acc[j][m][v], 0.0f) : acc[j][m][v];
                        hp[j][m][v] = fmaxf(r, __shfl_xor(r, 1));
                    }
            floatx4* xch = reinterpret_cast<floatx4*>(smem);
            const uint32_t slot = (wv >> 1) * 64u + lane;
            __syncthreads();  // every wave has finished reading the tile
            if (wv & 1u) {
#pragma unroll
                for (int j = 0; j < NJ; ++j)
#pragma unroll
                    for (int m = 0; m < 4; ++m) xch[(j * 4 + m) * 128 + slot] = hp[j][m];
            }
            __syncthreads();
            const int yr = y0 + static_cast<int>(wv);  // H is even: row yr + 1 exists when yr does
            if (!(wv & 1u) && yr < H && !(li & 1u)) {
                uint16_t* pb = a.pool + (img * (plane / 4u) + static_cast<size_t>(yr >> 1) * (a.W >> 1)) * a.cout + n0;
#pragma unroll
                for (int m = 0; m < 4; ++m) {
                    const uint32_t x = static_cast<uint32_t>(x0) + 16u * m + li;
                    if (x >= a.W) continue;
#pragma unroll
                    for (int j = 0; j < NJ; ++j) {
                        const floatx4 o = xch[(j * 4 + m) * 128 + slot];
                        *reinterpret_cast<uint2*>(pb + (x >> 1) * a.cout + 16u * j + 4u * lg) =
                            make_uint2(pack2(fmaxf(hp[j][m][0], o[0]), fmaxf(hp[j][m][1], o[1])),
                                       pack2(fmaxf(hp[j][m][2], o[2]), fmaxf(hp[j][m][3], o[3])));
                    }
                }
            }
        }
    }
    const int y = y0 + static_cast<int>(wv);
    if (y >= H) return;
    if constexpr (MODE == kModeRadOut) {
        // Output channels 0..2 (lanes 0..15 of the first channel tile):
        // radiance.
        if (n0 != 0 || lg != 0u || y >= static_cast<int>(a.fh)) return;
        float* ob = a.rad_out + static_cast<size_t>(y) * a.fw * 3u;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint32_t x = static_cast<uint32_t>(x0) + 16u * m + li;
            if (x >= a.fw) continue;
            float* o = ob + 3u * x;
#pragma unroll
            for (int v = 0; v < 3; ++v) o[v] = fmaxf(expm1f(acc[0][m][v]), 0.0f);
        }
    } else {
        if (a.relu) {
#pragma unroll
            for (int j = 0; j < NJ; ++j)
#pragma unroll
                for (int m = 0; m < 4; ++m)
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[j][m][v] = fmaxf(acc[j][m][v], 0.0f);
        }
        if constexpr (MODE == kModeMasked) {
            // Zero where the ReLU's output is not above zero (BF16 bits as a
            // signed 16-bit integer > 0).
            const uint16_t* mb =
                a.mask + (img * plane + static_cast<size_t>(y) * a.W) * a.mask_stride + a.mask_off + n0;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const uint32_t x = static_cast<uint32_t>(x0) + 16u * m + li;
                if (x >= a.W) continue;
                const uint32_t o = x * a.mask_stride + 4u * lg;
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    const uint2 mv = *reinterpret_cast<const uint2*>(mb + o + 16u * j);
                    const uint32_t w4[2] = {mv.x, mv.y};
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int16_t bits = static_cast<int16_t>(w4[v >> 1] >> (16 * (v & 1)));
                        if (bits <= 0) acc[j][m][v] = 0.0f;
                    }
                }
            }
        }
        uint16_t* yb = a.y + (img * plane + static_cast<size_t>(y) * a.W) * a.y_stride + a.y_off + n0;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const uint32_t x = static_cast<uint32_t>(x0) + 16u * m + li;
            if (x >= a.W) continue;
            const uint32_t o = x * a.y_stride + 4u * lg;
#pragma unroll
            for (int j = 0; j < NJ; ++j)
                *reinterpret_cast<uint2*>(yb + o + 16u * j) =
                    make_uint2(pack2(acc[j][m][0], acc[j][m][1]), pack2(acc[j][m][2], acc[j][m][3]));
        }
    }
}

// Canonical [9][cout][cin] weights into fragment order, one element per
// thread.
__global__ void k_frag_weights(const uint16_t* __restrict__ w, uint32_t cin, uint32_t cout, uint16_t* __restrict__ wf) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t per_tap = static_cast<size_t>(cout) * cin;
    if (i >= 9u * per_tap) return;
    const uint32_t tap = static_cast<uint32_t>(i / per_tap);
    const uint32_t r = static_cast<uint32_t>(i % per_tap);
    wf[frag_index(tap, r / cin, r % cin, cin, cout)] = w[i];
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

template <int NT, int MODE>
void conv_launch(const ConvArgs& a, dim3 grid, hipStream_t stream) {
    k_conv3x3<NT, MODE><<<grid, 256, 0, stream>>>(a);
}

}  // namespace

hipError_t launch_conv3x3(const ConvArgs& a, uint32_t min_blocks, hipStream_t stream) {
    if (a.cin == 0 || a.cin % 32u || a.cout % 32u || a.cu % 32u || a.cu > a.cin || a.x_off % 8u || a.x_stride % 8u)
        return hipErrorInvalidValue;
    const uint32_t batch = a.batch ? a.batch : 1u;
    if (batch > 1u && (a.rad || a.rad_out)) return hipErrorInvalidValue;
    if (a.mask && (a.rad_out || a.relu || a.mask_off % 4u || a.mask_stride % 4u || a.mask_off + a.cout > a.mask_stride ||
                   (reinterpret_cast<uintptr_t>(a.mask) & 7u)))
        return hipErrorInvalidValue;
    if (a.rad ? (a.cin != 32u || a.cu != 0u || !a.aov || a.fw > a.W || a.fh > a.H)
              : (a.cin > a.cu && (!a.x || a.x_off + (a.cin - a.cu) > a.x_stride)))
        return hipErrorInvalidValue;
    if (a.cu && (a.H % 2u || a.W % 2u || !a.xu)) return hipErrorInvalidValue;
    if (a.pool && (a.rad || a.rad_out || a.mask || a.H % 2u || a.W % 2u ||
                   (reinterpret_cast<uintptr_t>(a.pool) & 7u)))
        return hipErrorInvalidValue;
    // 16-byte loads of inputs and weights, 8-byte stores of outputs.
    const auto misaligned = [](const void* p, uintptr_t n) { return (reinterpret_cast<uintptr_t>(p) & (n - 1)) != 0; };
    if (!a.wf || misaligned(a.x, 16) || misaligned(a.xu, 16) || misaligned(a.wf, 16) || misaligned(a.y, 8) ||
        misaligned(a.bias, 4))
        return hipErrorInvalidValue;
    if (a.rad_out ? (a.fw > a.W || a.fh > a.H)
                  : (!a.y || a.y_off % 4u || a.y_stride % 4u || a.y_off + a.cout > a.y_stride))
        return hipErrorInvalidValue;
    const uint32_t tiles = ((a.W + kTileW - 1) / kTileW) * ((a.H + kTileRows - 1) / kTileRows) * batch;
    const uint32_t nt = (a.cout % 64u == 0 && tiles * (a.cout / 64u) >= min_blocks) ? 64u : 32u;
    if (static_cast<uint64_t>(a.cout / nt) * batch > 65535u) return hipErrorInvalidValue;
    dim3 grid((a.W + kTileW - 1) / kTileW, (a.H + kTileRows - 1) / kTileRows, a.cout / nt * batch);
    if (a.rad)
        conv_launch<32, kModePackedIn>(a, dim3(grid.x, grid.y, a.cout / 32u), stream);
    else if (a.rad_out)
        conv_launch<32, kModeRadOut>(a, dim3(grid.x, grid.y, a.cout / 32u), stream);
    else if (a.mask && nt == 64u)
        conv_launch<64, kModeMasked>(a, grid, stream);
    else if (a.mask)
        conv_launch<32, kModeMasked>(a, grid, stream);
    else if (nt == 64u)
        conv_launch<64, kModeTensor>(a, grid, stream);
    else
        conv_launch<32, kModeTensor>(a, grid, stream);
    return hipGetLastError();
}

hipError_t launch_frag_weights(const uint16_t* w, uint32_t cin, uint32_t cout, uint16_t* wf, hipStream_t stream) {
    if (!w || !wf || cin == 0 || cout == 0 || cin % 32u || cout % 32u) return hipErrorInvalidValue;
    const size_t n = 9ull * cout * cin;
    k_frag_weights<<<static_cast<uint32_t>((n + 255) / 256), 256, 0, stream>>>(w, cin, cout, wf);
    return hipGetLastError();
}

hipError_t launch_pool2(const uint16_t* x, uint32_t c, uint32_t H, uint32_t W, uint16_t* y, hipStream_t stream) {
    if (c % 8u || H % 2u || W % 2u) return hipErrorInvalidValue;
    const size_t n = static_cast<size_t>(H / 2u) * (W / 2u) * (c / 8u);
    k_pool2<<<static_cast<uint32_t>((n + 255) / 256), 256, 0, stream>>>(x, c, H, W, y);
    return hipGetLastError();
}

}  // namespace mrt
