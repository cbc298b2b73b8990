// C ABI of the learned denoiser: weight conversion (FP32 [cout][cin][3][3] to
// BF16 [tap][cout][cin], channels padded to multiples of 32), the activation
// arena of a frame and the layer sequence of the U-Net.
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "mrt_denoise.hpp"
#include "mrt_denoise_net.hpp"
#include "mrt_internal.hpp"

using mrt::DevBuf;
using mrt::fail;

namespace {

using mrt::unet::DevLayer;
using mrt::unet::kPad;
using mrt::unet::LayerShape;
using mrt::unet::Net;
using mrt::unet::resolve;
using mrt::unet::weight_count;

// Byte offsets of a frame's activations (H x W padded frame): per level l,
// t = first conv's output, s = encoder output (the skip), p = pooled input
// (l >= 1), o = decoder output; e_last = the coarsest level's output.
struct Arena {
    size_t total = 0;
    size_t t[MRT_DENOISER_MAX_LEVELS] = {}, s[MRT_DENOISER_MAX_LEVELS] = {}, p[MRT_DENOISER_MAX_LEVELS] = {},
           o[MRT_DENOISER_MAX_LEVELS] = {}, e_last = 0;
};

Arena plan(const Net& n, uint32_t H, uint32_t W) {
    Arena a;
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t r = at;
        at += (bytes + 255) / 256 * 256;
        return r;
    };
    const uint32_t L = n.levels;
    for (uint32_t l = 0; l < L; ++l) {
        const size_t px = static_cast<size_t>(H >> l) * (W >> l);
        a.t[l] = take(px * n.c[l] * 2);
        if (l + 1 < L) {
            a.s[l] = take(px * n.c[l] * 2);
            a.o[l] = take(px * n.c[l] * 2);
        }
        if (l > 0) a.p[l] = take(px * n.c[l - 1] * 2);
    }
    a.e_last = take(static_cast<size_t>(H >> (L - 1)) * (W >> (L - 1)) * n.c[L - 1] * 2);
    a.total = at;
    return a;
}

}  // namespace

struct mrt_denoiser {
    mrt_ctx* ctx = nullptr;
    Net net;
    std::vector<DevLayer> layers;
    DevBuf weights_f, biases, act, stage_rad, stage_aov, stage_out;  // weights_f: fragment order
    hipEvent_t ev[2] = {};
    float last_ms = 0.0f;
    ~mrt_denoiser() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (ctx) --ctx->children;
    }
};

namespace {

// One layer: input x (x_channels per pixel) or, with xu, the concatenation
// [up2(xu) (cu channels), x]; output y (layer's cout channels per pixel).
mrt::ConvArgs layer_args(const mrt_denoiser* dn, size_t layer, const uint16_t* x, uint32_t x_channels,
                         const uint16_t* xu, uint32_t cu, uint16_t* y, uint32_t H, uint32_t W) {
    const DevLayer& L = dn->layers[layer];
    mrt::ConvArgs a{};
    a.x = x;
    a.x_stride = x_channels;
    a.cin = L.cin;
    a.xu = xu;
    a.cu = cu;
    a.wf = dn->weights_f.as<uint16_t>() + L.w_off;
    a.bias = dn->biases.as<float>() + L.b_off;
    a.cout = L.cout;
    a.y = y;
    a.y_stride = L.cout;
    a.H = H;
    a.W = W;
    a.relu = 1;
    return a;
}

int run_net(mrt_denoiser* dn, const float* rad, float scale, const float* aov, uint32_t w, uint32_t h, float* out,
            hipStream_t st) {
    const Net& n = dn->net;
    const uint32_t L = n.levels, m = (1u << (L - 1)) - 1u;
    const uint32_t H = (h + m) & ~m, W = (w + m) & ~m;
    const uint32_t min_blocks = 2u * dn->ctx->num_cu;
    const Arena a = plan(n, H, W);
    int rc;
    if ((rc = dn->act.alloc(a.total)) != MRT_OK) return rc;
    char* base = dn->act.as<char>();
    auto bf = [&](size_t off) { return reinterpret_cast<uint16_t*>(base + off); };
    MRT_HIP(hipEventRecord(dn->ev[0], st));
    size_t layer = 0;
    for (uint32_t l = 0; l < L; ++l) {
        const uint32_t Hl = H >> l, Wl = W >> l;
        mrt::ConvArgs c = layer_args(dn, layer++, l == 0 ? nullptr : bf(a.p[l]), l == 0 ? 0u : n.c[l - 1], nullptr,
                                     0, bf(a.t[l]), Hl, Wl);
        if (l == 0) {
            c.rad = rad;
            c.aov = aov;
            c.rad_scale = scale;
            c.depth_scale = n.depth_scale;
            c.fw = w;
            c.fh = h;
        }
        MRT_HIP(mrt::launch_conv3x3(c, min_blocks, st));
        uint16_t* e = l + 1 < L ? bf(a.s[l]) : bf(a.e_last);
        mrt::ConvArgs cb = layer_args(dn, layer++, bf(a.t[l]), n.c[l], nullptr, 0, e, Hl, Wl);
        if (l + 1 < L) cb.pool = bf(a.p[l + 1]);  // the next level's input, pooled in the epilogue
        MRT_HIP(mrt::launch_conv3x3(cb, min_blocks, st));
    }
    for (int l = static_cast<int>(L) - 2; l >= 0; --l) {
        const uint32_t Hl = H >> l, Wl = W >> l;
        const uint16_t* coarse = l == static_cast<int>(L) - 2 ? bf(a.e_last) : bf(a.o[l + 1]);
        MRT_HIP(mrt::launch_conv3x3(layer_args(dn, layer++, bf(a.s[l]), n.c[l], coarse, n.c[l + 1], bf(a.t[l]), Hl, Wl),
                                    min_blocks, st));
        MRT_HIP(mrt::launch_conv3x3(layer_args(dn, layer++, bf(a.t[l]), n.c[l], nullptr, 0, bf(a.o[l]), Hl, Wl),
                                    min_blocks, st));
    }
    mrt::ConvArgs c = layer_args(dn, layer++, L == 1 ? bf(a.e_last) : bf(a.o[0]), n.c[0], nullptr, 0, nullptr, H, W);
    c.relu = 0;
    c.rad_out = out;
    c.fw = w;
    c.fh = h;
    MRT_HIP(mrt::launch_conv3x3(c, min_blocks, st));
    MRT_HIP(hipEventRecord(dn->ev[1], st));
    return MRT_OK;
}

}  // namespace

extern "C" {

MRT_API int mrt_denoiser_weight_count(const mrt_denoiser_desc* desc, uint64_t* count) {
    if (!count) return fail(MRT_E_INVALID, "null argument");
    Net n;
    int rc;
    if ((rc = resolve(desc, &n)) != MRT_OK) return rc;
    *count = weight_count(n);
    return MRT_OK;
}

MRT_API int mrt_denoiser_create(mrt_ctx* ctx, const mrt_denoiser_desc* desc, const float* weights, uint64_t count,
                                mrt_denoiser** out) {
    if (!ctx || !weights || !out) return fail(MRT_E_INVALID, "null argument");
    *out = nullptr;
    Net n;
    int rc;
    if ((rc = resolve(desc, &n)) != MRT_OK) return rc;
    if (count != weight_count(n))
        return fail(MRT_E_INVALID, "weight count " + std::to_string(count) + " does not match the network's " +
                                       std::to_string(weight_count(n)));
    for (uint64_t i = 0; i < count; ++i)
        if (!std::isfinite(weights[i])) return fail(MRT_E_INVALID, "non-finite denoiser weight");
    size_t n_w = 0, n_b = 0;
    std::vector<DevLayer> layers = mrt::unet::device_layers(n, &n_w, &n_b);
    std::vector<float> wf(n_w, 0.0f), bdev(n_b, 0.0f);
    mrt::unet::to_device_layout(layers, weights, wf.data(), bdev.data());
    // BF16 weights in fragment order, per layer.
    std::vector<uint16_t> wdev(n_w);
    for (const DevLayer& L : layers) {
        const size_t per_tap = static_cast<size_t>(L.cout) * L.cin;
        for (size_t i = 0; i < 9 * per_tap; ++i) {
            const uint32_t tap = static_cast<uint32_t>(i / per_tap), r = static_cast<uint32_t>(i % per_tap);
            wdev[L.w_off + mrt::frag_index(tap, r / L.cin, r % L.cin, L.cin, L.cout)] =
                mrt::unet::bf16_rne(wf[L.w_off + i]);
        }
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    MRT_HIP(hipSetDevice(ctx->device));
    auto dn = std::make_unique<mrt_denoiser>();
    dn->net = n;
    dn->layers = std::move(layers);
    for (hipEvent_t& e : dn->ev) MRT_HIP(hipEventCreate(&e));
    if ((rc = dn->weights_f.upload(wdev.data(), wdev.size() * 2, ctx->stream)) != MRT_OK) return rc;
    if ((rc = dn->biases.upload(bdev.data(), bdev.size() * 4, ctx->stream)) != MRT_OK) return rc;
    MRT_HIP(hipStreamSynchronize(ctx->stream));
    dn->ctx = ctx;
    ++ctx->children;
    *out = dn.release();
    return MRT_OK;
}

MRT_API int mrt_denoiser_run(mrt_denoiser* dn, const float* radiance, float radiance_scale, const float* aov,
                             uint32_t width, uint32_t height, float* out, int mem) {
    if (!dn || !radiance || !aov || !out) return fail(MRT_E_INVALID, "null argument");
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return fail(MRT_E_INVALID, "frame size must be 1..16384 per side");
    if (!std::isfinite(radiance_scale)) return fail(MRT_E_INVALID, "radiance_scale must be finite");
    mrt_ctx* ctx = dn->ctx;
    std::lock_guard<std::mutex> lock(ctx->mu);
    MRT_HIP(hipSetDevice(ctx->device));
    const hipStream_t st = ctx->stream;
    const size_t px = static_cast<size_t>(width) * height;
    const float *drad = radiance, *daov = aov;
    float* dout = out;
    int rc;
    if (mem != MRT_MEM_DEVICE) {
        if ((rc = dn->stage_rad.upload(radiance, px * 12, st)) != MRT_OK) return rc;
        if ((rc = dn->stage_aov.upload(aov, px * 28, st)) != MRT_OK) return rc;
        if ((rc = dn->stage_out.alloc(px * 12)) != MRT_OK) return rc;
        drad = dn->stage_rad.as<float>();
        daov = dn->stage_aov.as<float>();
        dout = dn->stage_out.as<float>();
    }
    if ((rc = run_net(dn, drad, radiance_scale, daov, width, height, dout, st)) != MRT_OK) return rc;
    if (mem != MRT_MEM_DEVICE) MRT_HIP(hipMemcpyAsync(out, dout, px * 12, hipMemcpyDeviceToHost, st));
    MRT_HIP(hipStreamSynchronize(st));
    MRT_HIP(hipEventElapsedTime(&dn->last_ms, dn->ev[0], dn->ev[1]));
    return MRT_OK;
}

MRT_API int mrt_denoiser_last_ms(const mrt_denoiser* dn, float* ms) {
    if (!dn || !ms) return fail(MRT_E_INVALID, "null argument");
    *ms = dn->last_ms;
    return MRT_OK;
}

MRT_API int mrt_denoiser_destroy(mrt_denoiser* dn) {
    if (!dn) return MRT_OK;
    {
        std::lock_guard<std::mutex> lock(dn->ctx->mu);
        (void)hipSetDevice(dn->ctx->device);
        (void)hipStreamSynchronize(dn->ctx->stream);
    }
    delete dn;
    return MRT_OK;
}

MRT_API int mrt_conv3x3_bf16(mrt_ctx* ctx, const uint16_t* x, uint32_t height, uint32_t width, uint32_t cin,
                             const uint16_t* w, const float* bias, uint32_t cout, int relu, uint16_t* y, int mem) {
    if (!ctx || !x || !w || !y) return fail(MRT_E_INVALID, "null argument");
    if (cin == 0 || cout == 0 || cin % kPad || cout % kPad || cin > 1024 || cout > 1024)
        return fail(MRT_E_INVALID, "cin and cout must be multiples of 32 up to 1024");
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return fail(MRT_E_INVALID, "frame size must be 1..16384 per side");
    if (mem == MRT_MEM_DEVICE &&
        ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(w)) % 16 || reinterpret_cast<uintptr_t>(y) % 8 ||
         reinterpret_cast<uintptr_t>(bias) % 4))
        return fail(MRT_E_INVALID, "device buffers must be aligned: x and w to 16 bytes, y to 8, bias to 4");
    std::lock_guard<std::mutex> lock(ctx->mu);
    MRT_HIP(hipSetDevice(ctx->device));
    const hipStream_t st = ctx->stream;
    const size_t px = static_cast<size_t>(width) * height;
    DevBuf bx, bw, bb, by, bwf;
    const uint16_t *dx = x, *dw = w;
    const float* db = bias;
    uint16_t* dy = y;
    int rc;
    if (mem != MRT_MEM_DEVICE) {
        if ((rc = bx.upload(x, px * cin * 2, st)) != MRT_OK) return rc;
        if ((rc = bw.upload(w, static_cast<size_t>(9) * cout * cin * 2, st)) != MRT_OK) return rc;
        if (bias && (rc = bb.upload(bias, cout * 4, st)) != MRT_OK) return rc;
        if ((rc = by.alloc(px * cout * 2)) != MRT_OK) return rc;
        dx = bx.as<uint16_t>();
        dw = bw.as<uint16_t>();
        db = bias ? bb.as<float>() : nullptr;
        dy = by.as<uint16_t>();
    }
    if ((rc = bwf.alloc(static_cast<size_t>(9) * cout * cin * 2)) != MRT_OK) return rc;
    MRT_HIP(mrt::launch_frag_weights(dw, cin, cout, bwf.as<uint16_t>(), st));
    mrt::ConvArgs a{};
    a.x = dx;
    a.x_stride = cin;
    a.cin = cin;
    a.wf = bwf.as<uint16_t>();
    a.bias = db;
    a.cout = cout;
    a.y = dy;
    a.y_stride = cout;
    a.H = height;
    a.W = width;
    a.relu = relu ? 1u : 0u;
    MRT_HIP(mrt::launch_conv3x3(a, 2u * ctx->num_cu, st));
    if (mem != MRT_MEM_DEVICE) MRT_HIP(hipMemcpyAsync(y, dy, px * cout * 2, hipMemcpyDeviceToHost, st));
    MRT_HIP(hipStreamSynchronize(st));
    return MRT_OK;
}

}  // extern "C"
