// Deterministic sampling for the GPU renderer.
//
// The reference draws come from host objects whose values depend on thread
// interleaving: the shared atomic counters of HaltonSeq/Stratified
// (app/Components/Samplers/HaltonSeq.cpp:14-22, Stratified.cpp:14-23) and the
// /dev/urandom-seeded thread_local mt19937 streams of MersenneTwister
// (MersenneTwister.cpp:9-15), Shader::getLight and getCosineSampleHemisphere
// (app/MobileRT/Shader.cpp:193-236).  Here every draw is a pure function of
// (seed, pixel, sample, dimension), so a pixel-sample's path is the same in a
// megakernel, a wavefront pipeline, or any launch order:
//   fmix32         MurmurHash3's 32-bit finaliser
//   key(s,p,n,d)   fmix32(fmix32(fmix32(fmix32(s + golden) ^ p) + n*K) ^ d)
//   u01(h)         (h >> 8) * 2^-24, exactly representable in [0, 1)
//   dimension      slot | light_sample << 4 | depth << 12 | branch << 16
// Sampler kinds (enum mrt_sampler_kind): CONSTANT returns its value; RANDOM
// u01(key); HALTON a base-2 radical inverse of the sample index XOR-scrambled
// per (pixel, dimension) for even slots and a base-3 radical inverse
// (haltonSequence, app/MobileRT/Utils.cpp:36-45) with a per-(pixel,
// dimension) rotation for odd slots; STRATIFIED one jittered stratum of spp
// per sample, rotated per (pixel, dimension).
#pragma once
#include "mrt_math.hpp"
#include "../../include/mobilert_gpu.h"

namespace mrt {

enum : uint32_t {
    kSlotJitterU = 0,
    kSlotJitterV = 1,
    kSlotLight = 2,
    kSlotLightR = 3,
    kSlotLightS = 4,
    kSlotRR = 5,
    kSlotHemi1 = 6,
    kSlotHemi2 = 7
};

MRT_HD uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

MRT_HD uint32_t sample_key(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t dim) {
    uint32_t h = fmix32(seed + 0x9e3779b9u);
    h = fmix32(h ^ pixel);
    h = fmix32(h + sample * 0x632be5abu);
    return fmix32(h ^ dim);
}

MRT_HD float unit_float(uint32_t h) { return static_cast<float>(h >> 8) * 5.9604644775390625e-08f; }

MRT_HD uint32_t draw_dim(uint32_t slot, uint32_t light_sample, uint32_t depth, uint32_t branch) {
    return slot | (light_sample << 4) | (depth << 12) | (branch << 16);
}

MRT_HD uint32_t bit_reverse(uint32_t v) { return __builtin_bitreverse32(v); }

// haltonSequence, app/MobileRT/Utils.cpp:36-45.
MRT_HD float halton(uint32_t index, uint32_t base) {
    float f = 1.0f;
    float result = 0.0f;
    const float fb = static_cast<float>(base);
    while (index > 0u) {
        f = f / fb;
        const float digit = static_cast<float>(index % base);
        result = result + f * digit;
        index = index / base;
    }
    return result;
}

MRT_HD float draw(const mrt_sampler& s, uint32_t pixel, uint32_t sample, uint32_t dim, uint32_t spp) {
    switch (s.kind) {
        case MRT_SAMPLER_CONSTANT:
            return s.value;
        case MRT_SAMPLER_RANDOM:
            return unit_float(sample_key(s.seed, pixel, sample, dim));
        case MRT_SAMPLER_HALTON: {
            const uint32_t scramble = sample_key(s.seed, pixel, 0xFFFFFFFFu, dim);
            if ((dim & 1u) == 0u) return unit_float(bit_reverse(sample) ^ scramble);
            float r = halton(sample, 3u) + unit_float(scramble);
            if (r >= 1.0f) r = r - 1.0f;
            return r;
        }
        case MRT_SAMPLER_STRATIFIED: {
            const uint32_t n = spp == 0u ? 1u : spp;
            const uint32_t rot = sample_key(s.seed, pixel, 0xFFFFFFFFu, dim) % n;
            const uint32_t stratum = (sample % n + rot) % n;
            const float j = unit_float(sample_key(s.seed, pixel, sample, dim));
            return (static_cast<float>(stratum) + j) / static_cast<float>(n);
        }
        default:
            return 0.0f;
    }
}

}  // namespace mrt
