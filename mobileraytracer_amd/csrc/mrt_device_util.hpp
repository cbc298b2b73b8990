// Launch-shape constants and small device helpers shared by the megakernel
// (mrt_kernels.hip) and the wavefront kernels (mrt_wavefront.hip).
#pragma once
#include "mrt_kernels.hpp"
#include "mrt_path.hpp"

namespace mrt {

constexpr uint32_t kBlock = 256;    // 4 wave64s per workgroup
constexpr uint32_t kStackLds = 16;  // traversal-stack entries per lane in LDS

__device__ __forceinline__ Stack make_stack(uint32_t* lds, const StackCfg& sc, uint32_t gtid) {
    Stack st;
    st.lds = (LdsWord*)(lds + threadIdx.x);  // address-space cast: needs the C form
    st.lds_stride = kBlock;
    st.spill = sc.spill ? sc.spill + gtid : nullptr;
    st.spill_stride = sc.spill_stride;
    st.cap_lds = kStackLds;
    st.cap_total = kStackLds + (sc.spill ? sc.cap_spill : 0u);
    st.sp = 0;
    st.overflow = 0;
    return st;
}

// Sum of a lane value over the wave, returned on every lane.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Counter slots: 0 camera, 1 extension, 2 shadow, 3 specular rays; 4..6 node
// pairs (closest, lights, any hit); 7..9 primitive tests (same order); 10/11
// lane iterations holding a ray / all lane iterations of the persistent trace.
template <bool COUNT>
__device__ __forceinline__ void flush_counts(const RayCounts& rc, unsigned long long* counters, uint32_t overflow,
                                             uint32_t* err) {
    uint32_t v[12] = {rc.camera, rc.extension, rc.shadow, rc.specular, 0, 0, 0, 0, 0, 0, 0, 0};
    if (COUNT) {
        for (int k = 0; k < 3; ++k) {
            v[4 + k] = rc.ts.pairs[k];
            v[7 + k] = rc.ts.prims[k];
        }
        v[10] = rc.ts.lane_iters;
        v[11] = rc.ts.lane_slots;
    }
    const uint32_t of = wave_sum(overflow);
    constexpr int kN = COUNT ? 12 : 4;
    uint32_t s[kN];
    for (int k = 0; k < kN; ++k) s[k] = wave_sum(v[k]);
    if ((threadIdx.x & 63u) == 0u) {
        for (int k = 0; k < kN; ++k)
            if (s[k]) atomicAdd(counters + k, static_cast<unsigned long long>(s[k]));
        if (of) atomicOr(err, 1u);
    }
}

// incrementalAvg, app/MobileRT/Utils.cpp:48-71.
__device__ __forceinline__ uint32_t incremental_avg(f3 c, uint32_t avg, uint32_t n) {
    const uint32_t lr = avg & 0xFFu, lg = (avg >> 8u) & 0xFFu, lb = (avg >> 16u) & 0xFFu;
    const uint32_t sr = static_cast<uint32_t>(c.x * 255.0f);
    const uint32_t sg = static_cast<uint32_t>(c.y * 255.0f);
    const uint32_t sb = static_cast<uint32_t>(c.z * 255.0f);
    const uint32_t cr = ((n - 1u) * lr + sr) / n;
    const uint32_t cg = ((n - 1u) * lg + sg) / n;
    const uint32_t cb = ((n - 1u) * lb + sb) / n;
    return 0xFF000000u | (min(cb, 255u) << 16u) | (min(cg, 255u) << 8u) | min(cr, 255u);
}

// Workgroup renumbering: XCD k (workgroups b with b % 8 == k) gets a
// contiguous range of logical workgroups.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
    const uint32_t xcd = b & 7u, idx = b >> 3;
    const uint32_t full = nb >> 3, rem = nb & 7u;
    return xcd < rem ? xcd * (full + 1u) + idx : rem * (full + 1u) + (xcd - rem) * full + idx;
}

// Pixel of covered-region slot `slot`: waves own 8x8 pixel blocks in
// row-major block order, lanes row-major inside the block.
__device__ __forceinline__ void slot_pixel(const RenderParams& p, uint32_t slot, uint32_t* x, uint32_t* y) {
    const uint32_t covW = 16u * p.bsx;
    const uint32_t tilesX = (covW + 7u) >> 3;
    const uint32_t wave = slot >> 6, lane = slot & 63u;
    *x = (wave % tilesX) * 8u + (lane & 7u);
    *y = (wave / tilesX) * 8u + (lane >> 3);
}

// Pixel (x, y) is rendered: inside the 16x16 tile grid (Renderer.cpp:20-25,
// 82-93) and its tile not skipped.
__device__ __forceinline__ bool pixel_covered(const RenderParams& p, uint32_t x, uint32_t y) {
    if (x >= 16u * p.bsx || y >= 16u * p.bsy) return false;
    const uint32_t t = (y / p.bsy) * 16u + x / p.bsx;
    return ((p.tile_skip[t >> 5] >> (t & 31u)) & 1u) == 0u;
}

// Pixel jitter draws of Renderer.cpp:96-99 for pixel (x, y), sample pass
// `sample`; *_COMPAT kinds replay the frame-wide counter in single-thread
// claim order (HaltonSeq.cpp:14-22, Stratified.cpp:14-23).
__device__ __forceinline__ void pixel_jitter(const RenderParams& p, uint32_t x, uint32_t y, uint32_t pid,
                                             uint32_t sample, float* r1, float* r2) {
    if (p.compat) {
        const uint32_t tx = x / p.bsx, ty = y / p.bsy;
        const uint32_t rank = p.tile_rank[ty * 16u + tx];
        const uint32_t ord = sample * p.pixels_per_pass + rank * p.bsx * p.bsy + (y - ty * p.bsy) * p.bsx +
                             (x - tx * p.bsx);
        const uint32_t c0 = 2u * ord, c1 = c0 + 1u;
        const uint32_t D = p.s_pixel.domain_size;
        const bool halton_kind = p.s_pixel.kind == MRT_SAMPLER_HALTON_COMPAT;
        *r1 = c0 < D ? (halton_kind ? halton(c0, 2u) : static_cast<float>(c0) / static_cast<float>(D)) : 1.0f;
        *r2 = c1 < D ? (halton_kind ? halton(c1, 2u) : static_cast<float>(c1) / static_cast<float>(D)) : 1.0f;
    } else {
        *r1 = draw(p.s_pixel, pid, sample, draw_dim(kSlotJitterU, 0u, 1u, 0u), p.spp);
        *r2 = draw(p.s_pixel, pid, sample, draw_dim(kSlotJitterV, 0u, 1u, 0u), p.spp);
    }
}

// Camera ray of pixel (x, y) in pass `sample` (Renderer.cpp:89-100).
__device__ __forceinline__ void primary_ray(const RenderParams& p, uint32_t x, uint32_t y, uint32_t pid,
                                            uint32_t sample, f3* o, f3* d) {
    float r1, r2;
    pixel_jitter(p, x, y, pid, sample, &r1, &r2);
    const float u = static_cast<float>(x) * p.inv_w;
    const float v = static_cast<float>(y) * p.inv_h;
    const float du = ((r1 - 0.5f) * 2.0f) * p.pix_w;
    const float dv = ((r2 - 0.5f) * 2.0f) * p.pix_h;
    camera_ray(p.cam, u, v, du, dv, o, d);
}

}  // namespace mrt
