// C ABI of the denoiser's trainer: the training pairs on the device, the
// crop schedule, the batched forward pass with every activation kept, the
// backward pass (backward-data convolutions, weight gradients, the gradients
// through pooling and upsampling) and Adam on FP32 master weights.
//
// Buffers of one step (B crops of side S; level l has S >> l pixels a side):
//   x0 (32 BF16 channels), t_l (first conv of encoder level l), e_l (encoder
//   output), p_l (pooled input of level l >= 1), h_l and d_l (decoder level
//   l's first conv and output), y (32 BF16 channels, 0..2 used), target;
//   backward: dy, three gradient tensors reused level after level, dcat_l
//   (the gradient of decoder level l's concatenated input, kept until the
//   encoder pass reaches level l), the octet copies and the weight-gradient
//   partial sums of the layer in hand.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "mrt_denoise.hpp"
#include "mrt_denoise_net.hpp"
#include "mrt_internal.hpp"

using mrt::DevBuf;
using mrt::fail;
using mrt::unet::DevLayer;
using mrt::unet::Net;

namespace {

constexpr uint32_t kMaxSplits = 1024;

uint64_t splitmix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct TrainArena {
    size_t total = 0;
    size_t x0 = 0, y = 0, tgt = 0, dy = 0, g[3] = {}, oct_x = 0, oct_dz = 0, part_w = 0, part_b = 0, loss = 0;
    size_t t[MRT_DENOISER_MAX_LEVELS] = {}, e[MRT_DENOISER_MAX_LEVELS] = {}, p[MRT_DENOISER_MAX_LEVELS] = {},
           h[MRT_DENOISER_MAX_LEVELS] = {}, d[MRT_DENOISER_MAX_LEVELS] = {}, dcat[MRT_DENOISER_MAX_LEVELS] = {};
};

}  // namespace

struct mrt_denoiser_trainer {
    mrt_ctx* ctx = nullptr;
    Net net;
    uint32_t batch = 8, crop = 256;
    double b1 = 0.9, b2 = 0.999, eps = 1e-8;
    uint64_t seed = 0;
    std::vector<DevLayer> layers;
    size_t n_w = 0, n_b = 0;
    uint32_t target_groups = 0;
    TrainArena ar;
    DevBuf params, grads, adam_m, adam_v, wqf, wtf, arena, data_x, data_t, stage_x, stage_t;
    uint32_t n_pairs = 0, fw = 0, fh = 0;
    uint64_t steps = 0;
    double last_loss = 0.0;
    float last_ms = 0.0f, ms_per_step = 0.0f;
    hipEvent_t ev[2] = {};
    ~mrt_denoiser_trainer() {
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        if (ctx) --ctx->children;
    }
};

namespace {

// Input of each layer as the octet transpose and the backward-data mask see
// it, for the activations of the arena at `base`.
struct LayerIo {
    mrt::TensorIn in;       // forward input
    uint16_t* out;          // forward output
    uint32_t level;         // frame side = crop >> level
    const uint16_t* mask;   // the input's ReLU output for the backward-data pass (null: none)
};

std::vector<LayerIo> layer_io(const mrt_denoiser_trainer* tr, char* base) {
    const Net& n = tr->net;
    const TrainArena& a = tr->ar;
    const uint32_t L = n.levels;
    auto bf = [&](size_t off) { return reinterpret_cast<uint16_t*>(base + off); };
    std::vector<LayerIo> io(tr->layers.size());
    auto plain = [](const uint16_t* x, uint32_t c) { return mrt::TensorIn{x, c, 0u, nullptr, 0u, c}; };
    for (uint32_t l = 0; l < L; ++l) {
        const uint32_t cin0 = tr->layers[2 * l].cin;
        io[2 * l] = {plain(l == 0 ? bf(a.x0) : bf(a.p[l]), cin0), bf(a.t[l]), l, nullptr};
        io[2 * l + 1] = {plain(bf(a.t[l]), n.c[l]), bf(a.e[l]), l, bf(a.t[l])};
    }
    for (int li = static_cast<int>(L) - 2; li >= 0; --li) {
        const uint32_t l = static_cast<uint32_t>(li);
        const uint32_t k = mrt::unet::decoder_layer(n, l);
        const uint16_t* coarse = l == L - 2 ? bf(a.e[L - 1]) : bf(a.d[l + 1]);
        io[k] = {mrt::TensorIn{bf(a.e[l]), n.c[l], 0u, coarse, n.c[l + 1], n.c[l + 1] + n.c[l]}, bf(a.h[l]), l,
                 nullptr};
        io[k + 1] = {plain(bf(a.h[l]), n.c[l]), bf(a.d[l]), l, bf(a.h[l])};
    }
    const uint16_t* last = L == 1 ? bf(a.e[0]) : bf(a.d[0]);
    io[4 * L - 2] = {plain(last, n.c[0]), bf(a.y), 0u, last};
    return io;
}

TrainArena plan_arena(const mrt_denoiser_trainer* tr) {
    const Net& n = tr->net;
    const uint32_t L = n.levels, B = tr->batch, S = tr->crop;
    TrainArena a;
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t r = at;
        at += (bytes + 255) / 256 * 256;
        return r;
    };
    auto px = [&](uint32_t l) { return static_cast<size_t>(B) * (S >> l) * (S >> l); };
    a.x0 = take(px(0) * 32 * 2);
    a.tgt = take(px(0) * 3 * 4);
    a.y = take(px(0) * 32 * 2);
    a.dy = take(px(0) * 32 * 2);
    size_t gmax = 0, octx = 0, octdz = 0, pw = 0, pb = 0;
    for (uint32_t l = 0; l < L; ++l) {
        a.t[l] = take(px(l) * n.c[l] * 2);
        a.e[l] = take(px(l) * n.c[l] * 2);
        if (l > 0) a.p[l] = take(px(l) * n.c[l - 1] * 2);
        if (l + 1 < L) {
            a.h[l] = take(px(l) * n.c[l] * 2);
            a.d[l] = take(px(l) * n.c[l] * 2);
            a.dcat[l] = take(px(l) * (n.c[l + 1] + n.c[l]) * 2);
        }
        gmax = std::max(gmax, px(l) * n.c[l] * 2);
        if (l > 0) gmax = std::max(gmax, px(l) * n.c[l - 1] * 2);
    }
    // Octet copies and partial sums: the largest layer.
    for (size_t k = 0; k < tr->layers.size(); ++k) {
        const DevLayer& d = tr->layers[k];
        uint32_t lvl;
        if (k < 2 * L)
            lvl = static_cast<uint32_t>(k / 2);
        else if (k + 1 < tr->layers.size())
            lvl = L - 2 - static_cast<uint32_t>((k - 2 * L) / 2);
        else
            lvl = 0;
        const uint32_t s = S >> lvl;
        octx = std::max(octx, static_cast<size_t>(B) * (s + 2) * s * d.cin * 2);
        octdz = std::max(octdz, px(lvl) * d.cout * 2);
        const uint32_t sp = mrt::wgrad_splits(B, s, s, d.cin, d.cout, kMaxSplits, tr->target_groups, nullptr);
        pw = std::max(pw, static_cast<size_t>(sp) * 9 * d.cout * d.cin * 4);
        pb = std::max(pb, static_cast<size_t>(sp) * d.cout * 4);
    }
    for (size_t& g : a.g) g = take(gmax);
    a.oct_x = take(octx);
    a.oct_dz = take(octdz);
    a.part_w = take(pw);
    a.part_b = take(pb);
    a.loss = take(sizeof(double));
    a.total = at;
    return a;
}

int adam_update(mrt_denoiser_trainer* tr, double lr, bool update, hipStream_t st) {
    mrt::AdamArgs a{};
    a.p = tr->params.as<float>();
    a.g = tr->grads.as<float>();
    a.m = tr->adam_m.as<float>();
    a.v = tr->adam_v.as<float>();
    a.wqf = tr->wqf.as<uint16_t>();
    a.wtf = tr->wtf.as<uint16_t>();
    a.n = tr->n_w + tr->n_b;
    a.n_w = tr->n_w;
    a.omb1 = static_cast<float>(1.0 - tr->b1);
    a.b2 = static_cast<float>(tr->b2);
    a.omb2 = static_cast<float>(1.0 - tr->b2);
    a.eps = static_cast<float>(tr->eps);
    a.update = update ? 1u : 0u;
    if (update) {
        const double t = static_cast<double>(tr->steps + 1);
        const double bc1 = 1.0 - std::pow(tr->b1, t);
        const double bc2 = 1.0 - std::pow(tr->b2, t);
        a.step_size = static_cast<float>(lr / bc1);
        a.bc2_sqrt = static_cast<float>(std::sqrt(bc2));
    }
    a.n_layers = static_cast<uint32_t>(tr->layers.size());
    for (uint32_t k = 0; k < a.n_layers; ++k)
        a.layer[k] = {static_cast<uint32_t>(tr->layers[k].w_off), tr->layers[k].cin, tr->layers[k].cout};
    MRT_HIP(mrt::launch_adam(a, st));
    if (update) ++tr->steps;
    return MRT_OK;
}

mrt::ConvArgs conv_args(const mrt_denoiser_trainer* tr, size_t k, const mrt::TensorIn& in, uint16_t* y,
                        uint32_t side, uint32_t batch) {
    const DevLayer& L = tr->layers[k];
    mrt::ConvArgs a{};
    a.x = in.x;
    a.x_stride = in.x_stride;
    a.x_off = in.x_off;
    a.cin = L.cin;
    a.xu = in.xu;
    a.cu = in.cu;
    a.wf = tr->wqf.as<uint16_t>() + L.w_off;
    a.bias = tr->params.as<float>() + tr->n_w + L.b_off;
    a.cout = L.cout;
    a.y = y;
    a.y_stride = L.cout;
    a.H = side;
    a.W = side;
    a.relu = k + 1 < tr->layers.size() ? 1u : 0u;
    a.batch = batch;
    return a;
}

// Forward pass over the batch in x0 (B crops), every activation kept.
int forward(mrt_denoiser_trainer* tr, uint32_t B, const std::vector<LayerIo>& io, char* base, hipStream_t st) {
    const Net& n = tr->net;
    const uint32_t L = n.levels, S = tr->crop, min_blocks = 2u * tr->ctx->num_cu;
    for (uint32_t l = 0; l < L; ++l) {
        const uint32_t s = S >> l;
        MRT_HIP(mrt::launch_conv3x3(conv_args(tr, 2 * l, io[2 * l].in, io[2 * l].out, s, B), min_blocks, st));
        MRT_HIP(mrt::launch_conv3x3(conv_args(tr, 2 * l + 1, io[2 * l + 1].in, io[2 * l + 1].out, s, B), min_blocks,
                                    st));
        if (l + 1 < L)
            MRT_HIP(mrt::launch_pool2(io[2 * l + 1].out, n.c[l], B * s, s,
                                      reinterpret_cast<uint16_t*>(base + tr->ar.p[l + 1]), st));
    }
    for (int li = static_cast<int>(L) - 2; li >= 0; --li) {
        const uint32_t l = static_cast<uint32_t>(li), s = S >> l;
        const uint32_t k = mrt::unet::decoder_layer(n, l);
        MRT_HIP(mrt::launch_conv3x3(conv_args(tr, k, io[k].in, io[k].out, s, B), min_blocks, st));
        MRT_HIP(mrt::launch_conv3x3(conv_args(tr, k + 1, io[k + 1].in, io[k + 1].out, s, B), min_blocks, st));
    }
    const size_t ko = 4 * L - 2;
    MRT_HIP(mrt::launch_conv3x3(conv_args(tr, ko, io[ko].in, io[ko].out, S, B), min_blocks, st));
    return MRT_OK;
}

// Weight and bias gradients of layer k from dz (the gradient at its output,
// cout channels) into grads.
int layer_wgrad(mrt_denoiser_trainer* tr, size_t k, const LayerIo& io, const uint16_t* dz, uint32_t B, char* base,
                hipStream_t st) {
    const DevLayer& d = tr->layers[k];
    const uint32_t s = tr->crop >> io.level;
    uint16_t* oct_x = reinterpret_cast<uint16_t*>(base + tr->ar.oct_x);
    uint16_t* oct_dz = reinterpret_cast<uint16_t*>(base + tr->ar.oct_dz);
    float* pw = reinterpret_cast<float*>(base + tr->ar.part_w);
    float* pb = reinterpret_cast<float*>(base + tr->ar.part_b);
    MRT_HIP(mrt::launch_octet(mrt::TensorIn{dz, d.cout, 0u, nullptr, 0u, d.cout}, B, s, s, 0, oct_dz, st));
    MRT_HIP(mrt::launch_octet(io.in, B, s, s, 1, oct_x, st));
    uint32_t splits = 0;
    MRT_HIP(mrt::launch_wgrad(oct_dz, oct_x, B, s, s, d.cin, d.cout, kMaxSplits, tr->target_groups, pw, pb, &splits, st));
    MRT_HIP(mrt::launch_sum_splits(pw, splits, static_cast<size_t>(9) * d.cout * d.cin, tr->grads.as<float>() + d.w_off,
                                   st));
    MRT_HIP(mrt::launch_sum_splits(pb, splits, d.cout, tr->grads.as<float>() + tr->n_w + d.b_off, st));
    return MRT_OK;
}

// Gradient at layer k's input from dz: the forward kernel over the
// transposed, tap-flipped weights, masked by the input's ReLU output when
// the input is one.
int layer_dgrad(mrt_denoiser_trainer* tr, size_t k, const LayerIo& io, const uint16_t* dz, uint16_t* out, uint32_t B,
                hipStream_t st) {
    const DevLayer& d = tr->layers[k];
    const uint32_t s = tr->crop >> io.level;
    mrt::ConvArgs a{};
    a.x = dz;
    a.x_stride = d.cout;
    a.cin = d.cout;
    a.wf = tr->wtf.as<uint16_t>() + d.w_off;
    a.cout = d.cin;
    a.y = out;
    a.y_stride = d.cin;
    a.H = s;
    a.W = s;
    a.batch = B;
    if (io.mask) {
        a.mask = io.mask;
        a.mask_stride = d.cin;
    }
    MRT_HIP(mrt::launch_conv3x3(a, 2u * tr->ctx->num_cu, st));
    return MRT_OK;
}

// Loss, then every gradient of the batch into grads.
int backward(mrt_denoiser_trainer* tr, uint32_t B, const std::vector<LayerIo>& io, char* base, hipStream_t st) {
    const Net& n = tr->net;
    const TrainArena& a = tr->ar;
    const uint32_t L = n.levels, S = tr->crop;
    auto bf = [&](size_t off) { return reinterpret_cast<uint16_t*>(base + off); };
    uint16_t* G[3] = {bf(a.g[0]), bf(a.g[1]), bf(a.g[2])};
    const size_t npx = static_cast<size_t>(B) * S * S;
    const float scale = static_cast<float>(2.0 / static_cast<double>(npx * 3));
    MRT_HIP(mrt::launch_mse(bf(a.y), reinterpret_cast<const float*>(base + a.tgt), npx, scale, bf(a.dy),
                            reinterpret_cast<double*>(base + a.loss), st));
    // Output conv: dz = dy; its input gradient goes to G[0].
    const size_t ko = 4 * L - 2;
    int rc;
    if ((rc = layer_wgrad(tr, ko, io[ko], bf(a.dy), B, base, st)) != MRT_OK) return rc;
    if ((rc = layer_dgrad(tr, ko, io[ko], bf(a.dy), G[0], B, st)) != MRT_OK) return rc;
    // Decoder levels 0 .. L-2: G[0] holds the gradient at d_l's conv output.
    for (uint32_t l = 0; l + 1 < L; ++l) {
        const uint32_t k = mrt::unet::decoder_layer(n, l), s = S >> l;
        if ((rc = layer_wgrad(tr, k + 1, io[k + 1], G[0], B, base, st)) != MRT_OK) return rc;
        if ((rc = layer_dgrad(tr, k + 1, io[k + 1], G[0], G[1], B, st)) != MRT_OK) return rc;
        if ((rc = layer_wgrad(tr, k, io[k], G[1], B, base, st)) != MRT_OK) return rc;
        if ((rc = layer_dgrad(tr, k, io[k], G[1], bf(a.dcat[l]), B, st)) != MRT_OK) return rc;
        // The upsampled channels' gradient, summed onto the coarse tensor.
        const uint16_t* coarse = io[k].in.xu;
        MRT_HIP(mrt::launch_up2_back(bf(a.dcat[l]), n.c[l + 1] + n.c[l], coarse, n.c[l + 1], B * s, s, G[0], st));
    }
    // Encoder levels L-1 .. 0: G[0] holds the gradient at e_l's conv output.
    for (int li = static_cast<int>(L) - 1; li >= 0; --li) {
        const uint32_t l = static_cast<uint32_t>(li), s = S >> l;
        if ((rc = layer_wgrad(tr, 2 * l + 1, io[2 * l + 1], G[0], B, base, st)) != MRT_OK) return rc;
        if ((rc = layer_dgrad(tr, 2 * l + 1, io[2 * l + 1], G[0], G[1], B, st)) != MRT_OK) return rc;
        if ((rc = layer_wgrad(tr, 2 * l, io[2 * l], G[1], B, base, st)) != MRT_OK) return rc;
        if (l == 0) break;
        // Gradient at the pooled input, routed to e_{l-1} with its skip part.
        if ((rc = layer_dgrad(tr, 2 * l, io[2 * l], G[1], G[2], B, st)) != MRT_OK) return rc;
        const uint32_t cs = n.c[l - 1];
        MRT_HIP(mrt::launch_pool2_back(bf(a.e[l - 1]), G[2], bf(a.dcat[l - 1]), n.c[l] + cs, n.c[l], cs, B * (2 * s),
                                       2 * s, G[0], st));
    }
    return MRT_OK;
}

int check_trainer(const mrt_denoiser_trainer* tr) {
    if (!tr) return fail(MRT_E_INVALID, "null argument");
    return MRT_OK;
}

// One step on the batch already in x0 / target: forward, backward, then
// Adam when lr > 0.
int run_step(mrt_denoiser_trainer* tr, uint32_t B, double lr, hipStream_t st) {
    char* base = tr->arena.as<char>();
    const std::vector<LayerIo> io = layer_io(tr, base);
    int rc;
    if ((rc = forward(tr, B, io, base, st)) != MRT_OK) return rc;
    if ((rc = backward(tr, B, io, base, st)) != MRT_OK) return rc;
    if (lr > 0.0 && (rc = adam_update(tr, lr, true, st)) != MRT_OK) return rc;
    return MRT_OK;
}

mrt::CropTable crop_table(const mrt_denoiser_trainer* tr, uint64_t step) {
    mrt::CropTable t{};
    for (uint32_t i = 0; i < tr->batch; ++i) {
        const uint64_t k = step * tr->batch + i;
        const uint64_t r1 = splitmix(tr->seed ^ splitmix(2 * k)), r2 = splitmix(tr->seed ^ splitmix(2 * k + 1));
        mrt::Crop& c = t.c[i];
        c.n = static_cast<uint32_t>(r1 % tr->n_pairs);
        c.oy = static_cast<uint32_t>((r1 >> 32) % (tr->fh - tr->crop + 1));
        c.ox = static_cast<uint32_t>(r2 % (tr->fw - tr->crop + 1));
        c.fx = static_cast<uint32_t>((r2 >> 40) & 1u);
        c.fy = static_cast<uint32_t>((r2 >> 41) & 1u);
    }
    return t;
}

double step_flops(const mrt_denoiser_trainer* tr) {
    const Net& n = tr->net;
    const std::vector<mrt::unet::LayerShape> shapes = mrt::unet::layer_shapes(n);
    const uint32_t L = n.levels;
    double f = 0.0;
    for (size_t k = 0; k < shapes.size(); ++k) {
        uint32_t lvl;
        if (k < 2 * L)
            lvl = static_cast<uint32_t>(k / 2);
        else if (k + 1 < shapes.size())
            lvl = L - 2 - static_cast<uint32_t>((k - 2 * L) / 2);
        else
            lvl = 0;
        const double px = static_cast<double>(tr->batch) * (tr->crop >> lvl) * (tr->crop >> lvl);
        const double one = 2.0 * 9.0 * shapes[k].cin * shapes[k].cout * px;
        f += one * (k == 0 ? 2.0 : 3.0);  // forward, weight gradient, backward-data (not the first layer)
    }
    return f;
}

}  // namespace

extern "C" {

MRT_API int mrt_denoiser_trainer_create(mrt_ctx* ctx, const mrt_denoiser_desc* net, const float* weights,
                                        uint64_t count, const mrt_denoiser_train_desc* desc,
                                        mrt_denoiser_trainer** out) {
    if (!ctx || !weights || !out) return fail(MRT_E_INVALID, "null argument");
    *out = nullptr;
    auto tr = std::make_unique<mrt_denoiser_trainer>();
    int rc;
    if ((rc = mrt::unet::resolve(net, &tr->net)) != MRT_OK) return rc;
    if (count != mrt::unet::weight_count(tr->net))
        return fail(MRT_E_INVALID, "weight count " + std::to_string(count) + " does not match the network's " +
                                       std::to_string(mrt::unet::weight_count(tr->net)));
    for (uint64_t i = 0; i < count; ++i)
        if (!std::isfinite(weights[i])) return fail(MRT_E_INVALID, "non-finite denoiser weight");
    if (desc) {
        if (desc->batch) tr->batch = desc->batch;
        if (desc->crop) tr->crop = desc->crop;
        if (desc->beta1 != 0.0) tr->b1 = desc->beta1;
        if (desc->beta2 != 0.0) tr->b2 = desc->beta2;
        if (desc->eps != 0.0) tr->eps = desc->eps;
        tr->seed = desc->seed;
    }
    if (tr->batch > mrt::kMaxCrops) return fail(MRT_E_INVALID, "batch must be 1..64");
    const uint32_t m = 8u << (tr->net.levels - 1);
    if (tr->crop % m || tr->crop > 4096)
        return fail(MRT_E_INVALID, "crop must be a multiple of " + std::to_string(m) + " up to 4096");
    if (!(tr->b1 >= 0.0 && tr->b1 < 1.0 && tr->b2 >= 0.0 && tr->b2 < 1.0 && tr->eps > 0.0 && std::isfinite(tr->eps)))
        return fail(MRT_E_INVALID, "Adam needs 0 <= beta < 1 and eps > 0");
    tr->layers = mrt::unet::device_layers(tr->net, &tr->n_w, &tr->n_b);
    tr->target_groups = 4u * ctx->num_cu;
    tr->ar = plan_arena(tr.get());
    std::vector<float> p(tr->n_w + tr->n_b, 0.0f);
    mrt::unet::to_device_layout(tr->layers, weights, p.data(), p.data() + tr->n_w);
    std::lock_guard<std::mutex> lock(ctx->mu);
    MRT_HIP(hipSetDevice(ctx->device));
    const hipStream_t st = ctx->stream;
    // The trainer holds one of the context's children from here on: its
    // destructor gives it back, also when creation fails below.
    tr->ctx = ctx;
    ++ctx->children;
    for (hipEvent_t& e : tr->ev) MRT_HIP(hipEventCreate(&e));
    const size_t np = p.size() * sizeof(float);
    if ((rc = tr->params.upload(p.data(), np, st)) != MRT_OK) return rc;
    if ((rc = tr->grads.alloc(np)) != MRT_OK || (rc = tr->adam_m.alloc(np)) != MRT_OK ||
        (rc = tr->adam_v.alloc(np)) != MRT_OK || (rc = tr->wqf.alloc(tr->n_w * 2)) != MRT_OK ||
        (rc = tr->wtf.alloc(tr->n_w * 2)) != MRT_OK || (rc = tr->arena.alloc(tr->ar.total)) != MRT_OK)
        return rc;
    MRT_HIP(hipMemsetAsync(tr->grads.p, 0, np, st));
    MRT_HIP(hipMemsetAsync(tr->adam_m.p, 0, np, st));
    MRT_HIP(hipMemsetAsync(tr->adam_v.p, 0, np, st));
    if ((rc = adam_update(tr.get(), 0.0, false, st)) != MRT_OK) return rc;
    MRT_HIP(hipStreamSynchronize(st));
    *out = tr.release();
    return MRT_OK;
}

MRT_API int mrt_denoiser_trainer_set_pairs(mrt_denoiser_trainer* tr, const float* inputs, const float* targets,
                                           uint32_t count, uint32_t width, uint32_t height, int mem) {
    int rc;
    if ((rc = check_trainer(tr)) != MRT_OK) return rc;
    if (!inputs || !targets || count == 0) return fail(MRT_E_INVALID, "null argument or no pairs");
    if (width < tr->crop || height < tr->crop || width > 16384 || height > 16384)
        return fail(MRT_E_INVALID, "frames must be crop..16384 pixels a side");
    const size_t px = static_cast<size_t>(count) * width * height;
    for (const float* v : {inputs, targets}) {
        if (mem != MRT_MEM_DEVICE) {
            const size_t nv = px * (v == inputs ? 10 : 3);
            for (size_t i = 0; i < nv; ++i)
                if (!std::isfinite(v[i])) return fail(MRT_E_INVALID, "non-finite training value");
        }
    }
    std::lock_guard<std::mutex> lock(tr->ctx->mu);
    MRT_HIP(hipSetDevice(tr->ctx->device));
    const hipStream_t st = tr->ctx->stream;
    if ((rc = tr->data_x.alloc(px * 40)) != MRT_OK || (rc = tr->data_t.alloc(px * 12)) != MRT_OK) return rc;
    const hipMemcpyKind kind = mem == MRT_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    MRT_HIP(hipMemcpyAsync(tr->data_x.p, inputs, px * 40, kind, st));
    MRT_HIP(hipMemcpyAsync(tr->data_t.p, targets, px * 12, kind, st));
    MRT_HIP(hipStreamSynchronize(st));
    tr->n_pairs = count;
    tr->fw = width;
    tr->fh = height;
    return MRT_OK;
}

MRT_API int mrt_denoiser_trainer_crops(const mrt_denoiser_trainer* tr, uint64_t step, uint32_t* crops) {
    int rc;
    if ((rc = check_trainer(tr)) != MRT_OK) return rc;
    if (!crops) return fail(MRT_E_INVALID, "null argument");
    if (!tr->n_pairs) return fail(MRT_E_INVALID, "no training pairs (mrt_denoiser_trainer_set_pairs)");
    const mrt::CropTable t = crop_table(tr, step);
    for (uint32_t i = 0; i < tr->batch; ++i) {
        const mrt::Crop& c = t.c[i];
        const uint32_t v[5] = {c.n, c.oy, c.ox, c.fx, c.fy};
        std::memcpy(crops + 5 * i, v, sizeof(v));
    }
    return MRT_OK;
}

MRT_API int mrt_denoiser_trainer_step(mrt_denoiser_trainer* tr, uint32_t steps, double lr, double* loss) {
    int rc;
    if ((rc = check_trainer(tr)) != MRT_OK) return rc;
    if (!tr->n_pairs) return fail(MRT_E_INVALID, "no training pairs (mrt_denoiser_trainer_set_pairs)");
    if (!(lr > 0.0) || !std::isfinite(lr)) return fail(MRT_E_INVALID, "lr must be positive");
    if (steps == 0) return fail(MRT_E_INVALID, "steps must be at least 1");
    std::lock_guard<std::mutex> lock(tr->ctx->mu);
    MRT_HIP(hipSetDevice(tr->ctx->device));
    const hipStream_t st = tr->ctx->stream;
    char* base = tr->arena.as<char>();
    MRT_HIP(hipMemsetAsync(base + tr->ar.loss, 0, sizeof(double), st));
    MRT_HIP(hipEventRecord(tr->ev[0], st));
    for (uint32_t i = 0; i < steps; ++i) {
        MRT_HIP(mrt::launch_gather(tr->data_x.as<float>(), tr->data_t.as<float>(), tr->fw, tr->fh,
                                   crop_table(tr, tr->steps), tr->batch, tr->crop,
                                   reinterpret_cast<uint16_t*>(base + tr->ar.x0),
                                   reinterpret_cast<float*>(base + tr->ar.tgt), st));
        if ((rc = run_step(tr, tr->batch, lr, st)) != MRT_OK) return rc;
    }
    MRT_HIP(hipEventRecord(tr->ev[1], st));
    double sum = 0.0;
    MRT_HIP(hipMemcpyAsync(&sum, base + tr->ar.loss, sizeof(double), hipMemcpyDeviceToHost, st));
    MRT_HIP(hipStreamSynchronize(st));
    MRT_HIP(hipEventElapsedTime(&tr->last_ms, tr->ev[0], tr->ev[1]));
    tr->ms_per_step = tr->last_ms / static_cast<float>(steps);
    const double n = static_cast<double>(tr->batch) * tr->crop * tr->crop * 3.0;
    tr->last_loss = sum / n / steps;
    if (loss) *loss = tr->last_loss;
    return MRT_OK;
}

MRT_API int mrt_denoiser_trainer_gradient(mrt_denoiser_trainer* tr, const float* inputs, const float* targets,
                                          uint32_t batch, double lr, double* loss, float* grad, uint64_t count,
                                          int mem) {
    int rc;
    if ((rc = check_trainer(tr)) != MRT_OK) return rc;
    if (!inputs || !targets) return fail(MRT_E_INVALID, "null argument");
    if (batch == 0 || batch > tr->batch) return fail(MRT_E_INVALID, "batch must be 1..the trainer's batch");
    if (grad && count != mrt::unet::weight_count(tr->net)) return fail(MRT_E_INVALID, "gradient count mismatch");
    if (!(lr >= 0.0) || !std::isfinite(lr)) return fail(MRT_E_INVALID, "lr must be finite and >= 0");
    const uint32_t S = tr->crop;
    const size_t px = static_cast<size_t>(batch) * S * S;
    std::lock_guard<std::mutex> lock(tr->ctx->mu);
    MRT_HIP(hipSetDevice(tr->ctx->device));
    const hipStream_t st = tr->ctx->stream;
    const float *dx = inputs, *dt = targets;
    if (mem != MRT_MEM_DEVICE) {
        if ((rc = tr->stage_x.upload(inputs, px * 40, st)) != MRT_OK) return rc;
        if ((rc = tr->stage_t.upload(targets, px * 12, st)) != MRT_OK) return rc;
        dx = tr->stage_x.as<float>();
        dt = tr->stage_t.as<float>();
    }
    // The batch as a data set of `batch` S x S frames, each taken whole, in
    // order.
    mrt::CropTable t{};
    for (uint32_t i = 0; i < batch; ++i) t.c[i] = {i, 0u, 0u, 0u, 0u};
    char* base = tr->arena.as<char>();
    MRT_HIP(hipMemsetAsync(base + tr->ar.loss, 0, sizeof(double), st));
    MRT_HIP(mrt::launch_gather(dx, dt, S, S, t, batch, S, reinterpret_cast<uint16_t*>(base + tr->ar.x0),
                               reinterpret_cast<float*>(base + tr->ar.tgt), st));
    if ((rc = run_step(tr, batch, 0.0, st)) != MRT_OK) return rc;
    std::vector<float> g;
    if (grad) g.resize(tr->n_w + tr->n_b);
    double sum = 0.0;
    MRT_HIP(hipMemcpyAsync(&sum, base + tr->ar.loss, sizeof(double), hipMemcpyDeviceToHost, st));
    if (grad) MRT_HIP(hipMemcpyAsync(g.data(), tr->grads.p, g.size() * 4, hipMemcpyDeviceToHost, st));
    if (lr > 0.0 && (rc = adam_update(tr, lr, true, st)) != MRT_OK) return rc;
    MRT_HIP(hipStreamSynchronize(st));
    if (loss) *loss = sum / (static_cast<double>(px) * 3.0);
    if (grad) mrt::unet::from_device_layout(tr->layers, g.data(), g.data() + tr->n_w, grad);
    return MRT_OK;
}

MRT_API int mrt_denoiser_trainer_weights(mrt_denoiser_trainer* tr, float* weights, uint64_t count) {
    int rc;
    if ((rc = check_trainer(tr)) != MRT_OK) return rc;
    if (!weights || count != mrt::unet::weight_count(tr->net)) return fail(MRT_E_INVALID, "weight count mismatch");
    std::vector<float> p(tr->n_w + tr->n_b);
    std::lock_guard<std::mutex> lock(tr->ctx->mu);
    MRT_HIP(hipSetDevice(tr->ctx->device));
    MRT_HIP(hipMemcpyAsync(p.data(), tr->params.p, p.size() * 4, hipMemcpyDeviceToHost, tr->ctx->stream));
    MRT_HIP(hipStreamSynchronize(tr->ctx->stream));
    mrt::unet::from_device_layout(tr->layers, p.data(), p.data() + tr->n_w, weights);
    return MRT_OK;
}

MRT_API int mrt_denoiser_trainer_stats(const mrt_denoiser_trainer* tr, mrt_denoiser_train_stats* stats) {
    int rc;
    if ((rc = check_trainer(tr)) != MRT_OK) return rc;
    if (!stats) return fail(MRT_E_INVALID, "null argument");
    *stats = mrt_denoiser_train_stats{};
    stats->steps = tr->steps;
    stats->last_loss = tr->last_loss;
    stats->last_ms = tr->last_ms;
    stats->ms_per_step = tr->ms_per_step;
    stats->flops_per_step = step_flops(tr);
    return MRT_OK;
}

MRT_API int mrt_denoiser_trainer_destroy(mrt_denoiser_trainer* tr) {
    if (!tr) return MRT_OK;
    {
        std::lock_guard<std::mutex> lock(tr->ctx->mu);
        (void)hipSetDevice(tr->ctx->device);
        (void)hipStreamSynchronize(tr->ctx->stream);
    }
    delete tr;
    return MRT_OK;
}

}  // extern "C"
