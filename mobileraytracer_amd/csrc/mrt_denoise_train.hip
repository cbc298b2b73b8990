// Training kernels of the denoiser's U-Net: weight gradients on MFMA (BF16
// operands, FP32 accumulation), the octet transposes that feed them, the
// loss, the gradients through pooling and upsampling, the crop gather and
// Adam.  The backward-data convolutions are the forward kernel
// (mrt_denoise.hip) over the transposed, tap-flipped weights with its masked
// output mode.
//
// Weight gradient = GEMM with the pixels as the reduction: dW[tap][co][ci] =
// sum_p dZ(p, co) X(p + offset(tap), ci).  An MFMA operand lane holds 8
// consecutive K values, here 8 consecutive pixels of one channel, which NHWC
// does not store together; the octet layout does (16 bytes per channel and
// 8-pixel segment), so both operands are single 16-byte loads.  A K step of
// v_mfma_f32_16x16x32_bf16 is 4 segments of 8 pixels (lane group l / 16 takes
// segment 4 k + l / 16), anywhere in the batch, so frames of any width that
// is a multiple of 8 fill every step.  The input's octet copy has a zero row
// above and below every frame, so a tap row is the same segment one row up or
// down; the column shifts dx = 0 and 2 are the segment's 16-byte operand
// shifted by one element in registers (v_alignbit), the element shifted in
// taken from the neighbouring segment of the same row (zero at the frame's
// edges).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mrt_denoise.hpp"

namespace mrt {

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pack2(float lo, float hi) {
    const bf16x2 b = __builtin_convertvector(floatx2{lo, hi}, bf16x2);
    uint32_t u;
    __builtin_memcpy(&u, &b, sizeof(u));
    return u;
}

__device__ __forceinline__ float bf16f(uint32_t bits16) { return __uint_as_float(bits16 << 16); }

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) {
    bf16x8 r;
    __builtin_memcpy(&r, &v, sizeof(r));
    return r;
}

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Element e (0..7) of 8 BF16 values held as a uint4.
__device__ __forceinline__ uint32_t elem(const uint4& q, int e) {
    const uint32_t w = (&q.x)[e >> 1];
    return (e & 1) ? (w >> 16) : (w & 0xFFFFu);
}

// ---------------------------------------------------------------- octet --
// Thread: one 8-channel group of one 8-pixel segment of one row.  It reads
// the segment's 8 pixels as 16-byte channel quads and writes 8 16-byte
// channel rows.
template <bool ROWPAD>
__global__ __launch_bounds__(256) void k_octet(TensorIn in, uint32_t batch, uint32_t H, uint32_t W, uint16_t* out) {
    const uint32_t W8 = W / 8u, cg = in.c / 8u;
    const uint32_t R = ROWPAD ? H + 2u : H;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t total = static_cast<size_t>(batch) * R * W8 * cg;
    if (i >= total) return;
    const uint32_t g = static_cast<uint32_t>(i % cg);
    const size_t seg = i / cg;  // (b R + r) W8 + xo
    const uint32_t xo = static_cast<uint32_t>(seg % W8);
    const size_t br = seg / W8;
    const uint32_t r = static_cast<uint32_t>(br % R), b = static_cast<uint32_t>(br / R);
    const int y = ROWPAD ? static_cast<int>(r) - 1 : static_cast<int>(r);
    constexpr int kPx = 8;
    const int xbase = static_cast<int>(xo * 8u);
    const uint32_t c0 = g * 8u;
    const bool up = c0 < in.cu;
    const size_t plane = static_cast<size_t>(H) * W;
    uint4 px[kPx];
#pragma unroll
    for (int k = 0; k < kPx; ++k) {
        const int x = xbase + k;
        px[k] = make_uint4(0u, 0u, 0u, 0u);
        if (y < 0 || y >= static_cast<int>(H) || x < 0 || x >= static_cast<int>(W)) continue;
        const uint16_t* src;
        if (up)
            src = in.xu + (b * (plane / 4u) + static_cast<size_t>(y >> 1) * (W >> 1) + (x >> 1)) * in.cu + c0;
        else
            src = in.x + (b * plane + static_cast<size_t>(y) * W + x) * in.x_stride + in.x_off + (c0 - in.cu);
        px[k] = *reinterpret_cast<const uint4*>(src);
    }
    uint16_t* o = out + (seg * in.c + c0) * 8u;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        // Channel c0 + e of pixels xo * 8 .. + 7.
        uint32_t w[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) w[j] = elem(px[2 * j], e) | (elem(px[2 * j + 1], e) << 16);
        *reinterpret_cast<uint4*>(o + e * 8) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// ---------------------------------------------------------------- wgrad --
// Workgroup task (co block of 32, ci block of 32, tap row dy, K split): its 4
// waves take the split's K steps in turn (wave w: steps k0 + w, k0 + w + 4,
// ...), each with 2 x 2 MFMA tiles for each of the 3 column shifts (12
// accumulators); the tasks of ci block 0 and dy = 1 also sum dZ against a
// ones operand (the bias gradient).  Lane (li = l % 16, lg = l / 16) loads,
// per K step, channel co0 + 16 jA + li of segment 4 k + lg (A) and channel
// ci0 + 16 jB + li of the same segment one row padded by dy (B); its results
// are C[co = 4 lg + v][ci = li] of each tile.  The waves' sums meet in LDS
// and are added in wave order (deterministic), one partial sum per split.
struct WgradArgs {
    const uint16_t* dzt;
    const uint16_t* xt;
    uint32_t batch, H, W8, cin, cout;
    uint32_t nseg, ksteps, kps, tasks;
    float* part_w;
    float* part_b;
};

constexpr int kWgradAcc = 56;  // 12 tiles x 4 + 2 bias tiles x 4 floats per lane

__global__ __launch_bounds__(256) void k_wgrad(WgradArgs a) {
    __shared__ float red[3][kWgradAcc][64];
    const uint32_t lane = threadIdx.x & 63u, wv = uniform(threadIdx.x >> 6);
    const uint32_t task = blockIdx.x;
    const uint32_t cob_n = a.cout / 32u, cib_n = a.cin / 32u;
    const uint32_t cob = task % cob_n;
    uint32_t rest = task / cob_n;
    const uint32_t cib = rest % cib_n;
    rest /= cib_n;
    const uint32_t dy = rest % 3u, split = rest / 3u;
    const uint32_t li = lane & 15u, lg = lane >> 4;
    const uint32_t co0 = cob * 32u, ci0 = cib * 32u;
    const bool bias = cib == 0u && dy == 1u;
    const uint32_t k0 = split * a.kps;
    const uint32_t k1 = min(a.ksteps, k0 + a.kps);
    const uint32_t frame_segs = a.H * a.W8;
    const size_t seg_el = static_cast<size_t>(a.cin) * 8u;  // elements per segment of the input copy
    floatx4 acc[kWgradAcc / 4];  // tile (d, i, j) at (d * 2 + i) * 2 + j, bias tiles at 12 + i
#pragma unroll
    for (int r = 0; r < kWgradAcc / 4; ++r) acc[r] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    // BF16 1.0 in all 8 K slots.
    const uint4 ones = make_uint4(0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u);
    const uint4 zero = make_uint4(0u, 0u, 0u, 0u);
    // Operands of one K step as loaded: dZ (A), the input segment (B for dx
    // = 1) and the neighbouring elements the shifted operands take.  The
    // next step's loads are issued before this step's MFMAs.
    struct Frag {
        uint4 av[2], c[2];
        uint32_t pv[2], nx[2];
    };
    auto load = [&](uint32_t ks, Frag& f) {
        const uint32_t s = ks * 4u + lg;
        const bool ok = ks < k1 && s < a.nseg;
        const uint32_t ss = ok ? s : 0u;
        const uint32_t b = ss / frame_segs;
        const uint32_t xo = ss % a.W8;
        const bool has_prev = ok && xo > 0u, has_next = ok && xo + 1u < a.W8;
        // Segment s in the row-padded copy, tap row dy.
        const size_t sx = ss + static_cast<size_t>(2u * b + dy) * a.W8;
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            // Loads under branches (a select between a global address and a
            // zero constant would become a flat load that may read private
            // memory).
            f.av[j] = zero;
            f.c[j] = zero;
            f.pv[j] = 0u;
            f.nx[j] = 0u;
            const uint16_t* xc = a.xt + sx * seg_el + (ci0 + 16u * j + li) * 8u;
            if (ok) {
                f.av[j] = *reinterpret_cast<const uint4*>(a.dzt + (static_cast<size_t>(ss) * a.cout + co0 + 16u * j + li) * 8u);
                f.c[j] = *reinterpret_cast<const uint4*>(xc);
            }
            // Last two elements of the previous segment, first two of the next.
            if (has_prev) f.pv[j] = *reinterpret_cast<const uint32_t*>(xc - seg_el + 6);
            if (has_next) f.nx[j] = *reinterpret_cast<const uint32_t*>(xc + seg_el);
        }
    };
    Frag cur, nxt;
    if (k0 + wv < k1) load(k0 + wv, cur);
    for (uint32_t ks = k0 + wv; ks < k1; ks += 4u) {
        load(ks + 4u, nxt);
        uint4 bv[3][2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const uint4 c = cur.c[j];
            // dx = 0: pixels x - 1; dx = 1: x; dx = 2: x + 1.
            bv[0][j] = make_uint4(__builtin_amdgcn_alignbit(c.x, cur.pv[j], 16), __builtin_amdgcn_alignbit(c.y, c.x, 16),
                                  __builtin_amdgcn_alignbit(c.z, c.y, 16), __builtin_amdgcn_alignbit(c.w, c.z, 16));
            bv[1][j] = c;
            bv[2][j] = make_uint4(__builtin_amdgcn_alignbit(c.y, c.x, 16), __builtin_amdgcn_alignbit(c.z, c.y, 16),
                                  __builtin_amdgcn_alignbit(c.w, c.z, 16), __builtin_amdgcn_alignbit(cur.nx[j], c.w, 16));
        }
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[(d * 2 + i) * 2 + j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        as_bf16x8(cur.av[i]), as_bf16x8(bv[d][j]), acc[(d * 2 + i) * 2 + j], 0, 0, 0);
        if (bias) {
#pragma unroll
            for (int i = 0; i < 2; ++i)
                acc[12 + i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(cur.av[i]), as_bf16x8(ones), acc[12 + i],
                                                                      0, 0, 0);
        }
        cur = nxt;
    }
    if (wv > 0u) {
#pragma unroll
        for (int r = 0; r < kWgradAcc / 4; ++r)
#pragma unroll
            for (int v = 0; v < 4; ++v) red[wv - 1u][4 * r + v][lane] = acc[r][v];
    }
    __syncthreads();
    if (wv != 0u) return;
#pragma unroll
    for (int r = 0; r < kWgradAcc / 4; ++r)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            acc[r][v] = ((acc[r][v] + red[0][4 * r + v][lane]) + red[1][4 * r + v][lane]) + red[2][4 * r + v][lane];
    const size_t layer = static_cast<size_t>(a.cout) * a.cin;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float* pw = a.part_w + (static_cast<size_t>(split) * 9u + dy * 3u + d) * layer;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    pw[static_cast<size_t>(co0 + 16u * i + 4u * lg + v) * a.cin + ci0 + 16u * j + li] =
                        acc[(d * 2 + i) * 2 + j][v];
    }
    if (bias && li == 0u) {
        float* pb = a.part_b + static_cast<size_t>(split) * a.cout + co0;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) pb[16u * i + 4u * lg + v] = acc[12 + i][v];
    }
}

// out[i] = sum over s of part[s n + i]: 64 outputs per workgroup, wave w
// summing splits w, w + 4, ... in order, the four sums added in wave order.
__global__ __launch_bounds__(256) void k_sum_splits(const float* __restrict__ part, uint32_t splits, size_t n,
                                                    float* __restrict__ out) {
    __shared__ float red[4][64];
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    const size_t i = static_cast<size_t>(blockIdx.x) * 64u + lane;
    float s = 0.0f;
    if (i < n) {
        uint32_t k = wv;
        for (; k + 12u < splits; k += 16u) {
            const float v0 = part[k * n + i], v1 = part[(k + 4u) * n + i], v2 = part[(k + 8u) * n + i],
                        v3 = part[(k + 12u) * n + i];
            s += v0;
            s += v1;
            s += v2;
            s += v3;
        }
        for (; k < splits; k += 4u) s += part[k * n + i];
    }
    red[wv][lane] = s;
    __syncthreads();
    if (wv == 0u && i < n) out[i] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// ----------------------------------------------------------------- loss --
__global__ __launch_bounds__(256) void k_mse(const uint16_t* __restrict__ y, const float* __restrict__ t, size_t n,
                                             float scale, uint16_t* __restrict__ dy, double* loss_sum) {
    __shared__ float red[256];
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    float sq = 0.0f;
    if (i < n) {
        const uint2 yv = *reinterpret_cast<const uint2*>(y + i * 32u);
        float d[3];
        d[0] = bf16f(yv.x & 0xFFFFu) - t[3 * i];
        d[1] = bf16f(yv.x >> 16) - t[3 * i + 1];
        d[2] = bf16f(yv.y & 0xFFFFu) - t[3 * i + 2];
        sq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        uint4* o = reinterpret_cast<uint4*>(dy + i * 32u);
        o[0] = make_uint4(pack2(d[0] * scale, d[1] * scale), pack2(d[2] * scale, 0.0f), 0u, 0u);
        o[1] = o[2] = o[3] = make_uint4(0u, 0u, 0u, 0u);
    }
    red[threadIdx.x] = sq;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss_sum, static_cast<double>(red[0]));
}

// ------------------------------------------------------- pool / upsample --
__device__ __forceinline__ bool positive(uint32_t bits16) { return static_cast<int16_t>(bits16) > 0; }

__global__ __launch_bounds__(256) void k_pool2_back(const uint16_t* __restrict__ e, const uint16_t* __restrict__ dp,
                                                    const uint16_t* __restrict__ dskip, uint32_t dskip_stride,
                                                    uint32_t dskip_off, uint32_t c, uint32_t n_rows, uint32_t W,
                                                    uint16_t* __restrict__ dz) {
    const uint32_t W2 = W / 2u, cq = c / 8u;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<size_t>(n_rows / 2u) * W2 * cq) return;
    const uint32_t q = static_cast<uint32_t>(i % cq);
    const size_t p = i / cq;
    const uint32_t px = static_cast<uint32_t>(p % W2);
    const size_t py = p / W2;
    size_t pix[4];
    uint4 ev[4], sv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        pix[k] = (2u * py + (k >> 1)) * W + 2u * px + (k & 1);
        ev[k] = *reinterpret_cast<const uint4*>(e + pix[k] * c + q * 8u);
        sv[k] = *reinterpret_cast<const uint4*>(dskip + pix[k] * dskip_stride + dskip_off + q * 8u);
    }
    const uint4 dv = *reinterpret_cast<const uint4*>(dp + p * c + q * 8u);
    uint32_t o[4][4];
#pragma unroll
    for (int ch = 0; ch < 8; ++ch) {
        float mx = bf16f(elem(ev[0], ch));
        int arg = 0;
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            const float v = bf16f(elem(ev[k], ch));
            if (v > mx) {
                mx = v;
                arg = k;
            }
        }
        const float g = bf16f(elem(dv, ch));
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float s = bf16f(elem(sv[k], ch));
            if (k == arg) s = s + g;
            if (!positive(elem(ev[k], ch))) s = 0.0f;
            const uint32_t b = pack2(s, 0.0f) & 0xFFFFu;
            if (ch & 1)
                o[k][ch >> 1] |= b << 16;
            else
                o[k][ch >> 1] = b;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k)
        *reinterpret_cast<uint4*>(dz + pix[k] * c + q * 8u) = make_uint4(o[k][0], o[k][1], o[k][2], o[k][3]);
}

__global__ __launch_bounds__(256) void k_up2_back(const uint16_t* __restrict__ dcat, uint32_t dcat_stride,
                                                  const uint16_t* __restrict__ coarse, uint32_t c, uint32_t n_rows,
                                                  uint32_t W, uint16_t* __restrict__ dz) {
    const uint32_t W2 = W / 2u, cq = c / 8u;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<size_t>(n_rows / 2u) * W2 * cq) return;
    const uint32_t q = static_cast<uint32_t>(i % cq);
    const size_t p = i / cq;
    const uint32_t px = static_cast<uint32_t>(p % W2);
    const size_t py = p / W2;
    uint4 gv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        gv[k] = *reinterpret_cast<const uint4*>(dcat + ((2u * py + (k >> 1)) * W + 2u * px + (k & 1)) * dcat_stride +
                                                q * 8u);
    const uint4 cv = *reinterpret_cast<const uint4*>(coarse + p * c + q * 8u);
    uint32_t o[4];
#pragma unroll
    for (int ch = 0; ch < 8; ++ch) {
        float s = bf16f(elem(gv[0], ch));
        s = s + bf16f(elem(gv[1], ch));
        s = s + bf16f(elem(gv[2], ch));
        s = s + bf16f(elem(gv[3], ch));
        if (!positive(elem(cv, ch))) s = 0.0f;
        const uint32_t b = pack2(s, 0.0f) & 0xFFFFu;
        if (ch & 1)
            o[ch >> 1] |= b << 16;
        else
            o[ch >> 1] = b;
    }
    *reinterpret_cast<uint4*>(dz + p * c + q * 8u) = make_uint4(o[0], o[1], o[2], o[3]);
}

// --------------------------------------------------------------- gather --
__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ xin, const float* __restrict__ tin,
                                                uint32_t fw, uint32_t fh, CropTable crops, uint32_t S,
                                                uint16_t* __restrict__ x, float* __restrict__ t) {
    const uint32_t b = blockIdx.y;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<size_t>(S) * S) return;
    const uint32_t yy = static_cast<uint32_t>(i / S), xx = static_cast<uint32_t>(i % S);
    const Crop c = crops.c[b];
    const uint32_t sy = c.oy + (c.fy ? S - 1u - yy : yy), sx = c.ox + (c.fx ? S - 1u - xx : xx);
    const size_t src = (static_cast<size_t>(c.n) * fh + sy) * fw + sx;
    const float* v = xin + src * 10u;
    const size_t dst = static_cast<size_t>(b) * S * S + i;
    uint4* o = reinterpret_cast<uint4*>(x + dst * 32u);
    o[0] = make_uint4(pack2(v[0], v[1]), pack2(v[2], v[3]), pack2(v[4], v[5]), pack2(v[6], v[7]));
    o[1] = make_uint4(pack2(v[8], v[9]), 0u, 0u, 0u);
    o[2] = o[3] = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
    for (int k = 0; k < 3; ++k) t[dst * 3u + k] = tin[src * 3u + k];
}

// ----------------------------------------------------------------- adam --
__global__ __launch_bounds__(256) void k_adam(AdamArgs a) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    float p = a.p[i];
    if (a.update) {
        const float g = a.g[i];
        float m = a.m[i], v = a.v[i];
        m = m + a.omb1 * (g - m);
        v = v * a.b2 + (a.omb2 * g) * g;
        const float denom = sqrtf(v) / a.bc2_sqrt + a.eps;
        p = p + (-a.step_size) * (m / denom);
        a.m[i] = m;
        a.v[i] = v;
        a.p[i] = p;
    }
    if (i >= a.n_w) return;
    const uint16_t q = static_cast<uint16_t>(pack2(p, 0.0f) & 0xFFFFu);
    uint32_t k = 0;
    while (k + 1 < a.n_layers && i >= a.layer[k + 1].w_off) ++k;
    const AdamLayer L = a.layer[k];
    const uint32_t local = static_cast<uint32_t>(i - L.w_off);
    const uint32_t per_tap = L.cout * L.cin;
    const uint32_t tap = local / per_tap, r = local % per_tap;
    const uint32_t co = r / L.cin, ci = r % L.cin;
    a.wqf[L.w_off + frag_index(tap, co, ci, L.cin, L.cout)] = q;
    a.wtf[L.w_off + frag_index(8u - tap, ci, co, L.cout, L.cin)] = q;
}

uint32_t blocks_for(size_t n) { return static_cast<uint32_t>((n + 255) / 256); }

}  // namespace

hipError_t launch_octet(const TensorIn& in, uint32_t batch, uint32_t H, uint32_t W, int rowpad, uint16_t* out,
                        hipStream_t stream) {
    if (W % 8u || in.c % 8u || in.cu % 8u || in.cu > in.c || in.x_stride % 8u || in.x_off % 8u || !out)
        return hipErrorInvalidValue;
    if ((in.cu && (!in.xu || H % 2u || W % 2u)) || (in.c > in.cu && (!in.x || in.x_off + in.c - in.cu > in.x_stride)))
        return hipErrorInvalidValue;
    const size_t n = static_cast<size_t>(batch) * (rowpad ? H + 2u : H) * (W / 8u) * (in.c / 8u);
    if (n == 0) return hipSuccess;
    if (rowpad)
        k_octet<true><<<blocks_for(n), 256, 0, stream>>>(in, batch, H, W, out);
    else
        k_octet<false><<<blocks_for(n), 256, 0, stream>>>(in, batch, H, W, out);
    return hipGetLastError();
}

hipError_t launch_wgrad(const uint16_t* dzt, const uint16_t* xt, uint32_t batch, uint32_t H, uint32_t W, uint32_t cin,
                        uint32_t cout, uint32_t max_splits, uint32_t target_groups, float* part_w, float* part_b,
                        uint32_t* splits, hipStream_t stream) {
    if (cin % 32u || cout % 32u || W % 8u || !batch || !H || !max_splits) return hipErrorInvalidValue;
    WgradArgs a{};
    a.dzt = dzt;
    a.xt = xt;
    a.batch = batch;
    a.H = H;
    a.W8 = W / 8u;
    a.cin = cin;
    a.cout = cout;
    a.nseg = batch * H * a.W8;
    a.ksteps = (a.nseg + 3u) / 4u;
    const uint32_t per_split = 3u * (cin / 32u) * (cout / 32u);
    const uint32_t s = wgrad_splits(batch, H, W, cin, cout, max_splits, target_groups, &a.kps);
    a.tasks = s * per_split;
    a.part_w = part_w;
    a.part_b = part_b;
    // Every split has the (co block, ci block 0, dy 1) task of each co block,
    // so every bias partial is written.
    k_wgrad<<<a.tasks, 256, 0, stream>>>(a);
    *splits = s;
    return hipGetLastError();
}

hipError_t launch_sum_splits(const float* part, uint32_t splits, size_t n, float* out, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    if (n > (static_cast<size_t>(1) << 36)) return hipErrorInvalidValue;
    k_sum_splits<<<static_cast<uint32_t>((n + 63) / 64), 256, 0, stream>>>(part, splits, n, out);
    return hipGetLastError();
}

hipError_t launch_mse(const uint16_t* y, const float* t, size_t n, float scale, uint16_t* dy, double* loss_sum,
                      hipStream_t stream) {
    if (n == 0) return hipSuccess;
    k_mse<<<blocks_for(n), 256, 0, stream>>>(y, t, n, scale, dy, loss_sum);
    return hipGetLastError();
}

hipError_t launch_pool2_back(const uint16_t* e, const uint16_t* dp, const uint16_t* dskip, uint32_t dskip_stride,
                             uint32_t dskip_off, uint32_t c, uint32_t n_rows, uint32_t W, uint16_t* dz,
                             hipStream_t stream) {
    if (c % 8u || n_rows % 2u || W % 2u || dskip_stride % 8u || dskip_off % 8u || dskip_off + c > dskip_stride)
        return hipErrorInvalidValue;
    const size_t n = static_cast<size_t>(n_rows / 2u) * (W / 2u) * (c / 8u);
    if (n == 0) return hipSuccess;
    k_pool2_back<<<blocks_for(n), 256, 0, stream>>>(e, dp, dskip, dskip_stride, dskip_off, c, n_rows, W, dz);
    return hipGetLastError();
}

hipError_t launch_up2_back(const uint16_t* dcat, uint32_t dcat_stride, const uint16_t* coarse, uint32_t c,
                           uint32_t n_rows, uint32_t W, uint16_t* dz, hipStream_t stream) {
    if (c % 8u || n_rows % 2u || W % 2u || dcat_stride % 8u || c > dcat_stride) return hipErrorInvalidValue;
    const size_t n = static_cast<size_t>(n_rows / 2u) * (W / 2u) * (c / 8u);
    if (n == 0) return hipSuccess;
    k_up2_back<<<blocks_for(n), 256, 0, stream>>>(dcat, dcat_stride, coarse, c, n_rows, W, dz);
    return hipGetLastError();
}

hipError_t launch_gather(const float* xin, const float* tin, uint32_t fw, uint32_t fh, const CropTable& crops,
                         uint32_t batch, uint32_t S, uint16_t* x, float* t, hipStream_t stream) {
    if (batch == 0 || batch > kMaxCrops || S == 0 || S > fw || S > fh) return hipErrorInvalidValue;
    for (uint32_t b = 0; b < batch; ++b)
        if (crops.c[b].oy + S > fh || crops.c[b].ox + S > fw) return hipErrorInvalidValue;
    k_gather<<<dim3(blocks_for(static_cast<size_t>(S) * S), batch), 256, 0, stream>>>(xin, tin, fw, fh, crops, S, x, t);
    return hipGetLastError();
}

hipError_t launch_adam(const AdamArgs& a, hipStream_t stream) {
    if (a.n_layers == 0 || a.n_layers > kMaxLayers || a.n_w > a.n || a.layer[0].w_off != 0) return hipErrorInvalidValue;
    if (a.n == 0) return hipSuccess;
    k_adam<<<blocks_for(a.n), 256, 0, stream>>>(a);
    return hipGetLastError();
}

}  // namespace mrt
