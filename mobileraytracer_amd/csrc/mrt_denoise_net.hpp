// The denoiser's U-Net as the host sees it: shape (levels, channels), layer
// order and sizes, the device weight layout (channels padded to multiples of
// 32) and BF16 rounding.  Shared by inference (mrt_denoise_api.cpp) and
// training (mrt_denoise_train_api.cpp).
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "mrt_internal.hpp"

namespace mrt {
namespace unet {

constexpr uint32_t kInChannels = 10, kOutChannels = 3, kPad = 32, kMaxChannels = 256;

inline uint16_t bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return static_cast<uint16_t>((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    return static_cast<uint16_t>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

struct Net {
    uint32_t levels = 5;
    uint32_t c[MRT_DENOISER_MAX_LEVELS] = {32, 64, 96, 128, 128};
    float depth_scale = 1.0f;
};

inline int resolve(const mrt_denoiser_desc* d, Net* n) {
    if (!d) return MRT_OK;
    if (d->levels > MRT_DENOISER_MAX_LEVELS) return fail(MRT_E_INVALID, "denoiser levels must be 1..5");
    if (d->levels) n->levels = d->levels;
    bool any = false;
    for (uint32_t l = 0; l < MRT_DENOISER_MAX_LEVELS; ++l) any = any || d->channels[l];
    if (any) {
        for (uint32_t l = 0; l < n->levels; ++l) {
            const uint32_t c = d->channels[l];
            if (c == 0 || c % kPad || c > kMaxChannels)
                return fail(MRT_E_INVALID, "denoiser channels must be multiples of 32 in 32..256");
            n->c[l] = c;
        }
    }
    if (d->depth_scale != 0.0f) {
        if (!std::isfinite(d->depth_scale)) return fail(MRT_E_INVALID, "depth_scale must be finite");
        n->depth_scale = d->depth_scale;
    }
    return MRT_OK;
}

struct LayerShape {
    uint32_t cin, cout;  // logical
};

// Layers in weight order: encoder 0..L-1 (two each), decoder L-2..0 (two
// each), output.
inline std::vector<LayerShape> layer_shapes(const Net& n) {
    std::vector<LayerShape> v;
    const uint32_t L = n.levels;
    for (uint32_t l = 0; l < L; ++l) {
        v.push_back({l == 0 ? kInChannels : n.c[l - 1], n.c[l]});
        v.push_back({n.c[l], n.c[l]});
    }
    for (int l = static_cast<int>(L) - 2; l >= 0; --l) {
        v.push_back({n.c[l + 1] + n.c[l], n.c[l]});
        v.push_back({n.c[l], n.c[l]});
    }
    v.push_back({n.c[0], kOutChannels});
    return v;
}

// Index of the first convolution of decoder level l (0 <= l <= L - 2); the
// second is the next one; the output convolution is 4 L - 2.
inline uint32_t decoder_layer(const Net& n, uint32_t l) { return 2u * n.levels + 2u * (n.levels - 2u - l); }

inline uint32_t pad32(uint32_t c) { return (c + kPad - 1) / kPad * kPad; }

inline uint64_t weight_count(const Net& n) {
    uint64_t k = 0;
    for (const LayerShape& s : layer_shapes(n)) k += static_cast<uint64_t>(s.cout) * s.cin * 9 + s.cout;
    return k;
}

struct DevLayer {
    uint32_t cin, cout;   // padded
    uint32_t lcin, lcout; // logical
    size_t w_off, b_off;  // elements into the weight / bias buffers
};

// Layout of every layer's parameters: weights [tap][cout][cin] (padded
// channels zero) one layer after the other (the convolutions read them in
// fragment order, mrt::frag_index, from the same offsets), FP32 biases
// [cout] likewise.
inline std::vector<DevLayer> device_layers(const Net& n, size_t* n_w, size_t* n_b) {
    std::vector<DevLayer> v;
    size_t w = 0, b = 0;
    for (const LayerShape& s : layer_shapes(n)) {
        DevLayer d{pad32(s.cin), pad32(s.cout), s.cin, s.cout, w, b};
        w += static_cast<size_t>(9) * d.cout * d.cin;
        b += d.cout;
        v.push_back(d);
    }
    *n_w = w;
    *n_b = b;
    return v;
}

// Flat weights ([cout][cin][3][3] then bias per layer) to the device layout,
// FP32 (w: n_w entries, b: n_b entries).
inline void to_device_layout(const std::vector<DevLayer>& layers, const float* flat, float* w, float* b) {
    const float* src = flat;
    for (const DevLayer& d : layers) {
        for (uint32_t o = 0; o < d.lcout; ++o)
            for (uint32_t i = 0; i < d.lcin; ++i)
                for (uint32_t tap = 0; tap < 9; ++tap)
                    w[d.w_off + (static_cast<size_t>(tap) * d.cout + o) * d.cin + i] =
                        src[(static_cast<size_t>(o) * d.lcin + i) * 9 + tap];
        src += static_cast<size_t>(d.lcout) * d.lcin * 9;
        for (uint32_t o = 0; o < d.lcout; ++o) b[d.b_off + o] = src[o];
        src += d.lcout;
    }
}

// The inverse: device layout (FP32) to flat weights.
inline void from_device_layout(const std::vector<DevLayer>& layers, const float* w, const float* b, float* flat) {
    float* dst = flat;
    for (const DevLayer& d : layers) {
        for (uint32_t o = 0; o < d.lcout; ++o)
            for (uint32_t i = 0; i < d.lcin; ++i)
                for (uint32_t tap = 0; tap < 9; ++tap)
                    dst[(static_cast<size_t>(o) * d.lcin + i) * 9 + tap] =
                        w[d.w_off + (static_cast<size_t>(tap) * d.cout + o) * d.cin + i];
        dst += static_cast<size_t>(d.lcout) * d.lcin * 9;
        for (uint32_t o = 0; o < d.lcout; ++o) dst[o] = b[d.b_off + o];
        dst += d.lcout;
    }
}

}  // namespace unet
}  // namespace mrt
