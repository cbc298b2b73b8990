// gfx950 kernels of the ray-scene path and their host launchers.
//
// Launch shape: 256-thread workgroups (4 wave64s).  Each lane traces its own
// ray; the per-lane traversal stack keeps its first kStackLds entries in LDS
// (16 KiB per workgroup) and spills deeper entries to a global area of
// cap_spill entries per lane, sized from the BVH depth by the host.
// Render passes map one wave to an 8x8 pixel block and renumber workgroups so
// that each of the 8 XCDs (which receive workgroups round-robin) works on a
// contiguous band of the image, keeping neighbouring rays in one L2.
#include "mrt_kernels.hpp"

#include "mrt_path.hpp"

namespace mrt {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kStackLds = 16;

__device__ __forceinline__ Stack make_stack(uint32_t* lds, const StackCfg& sc, uint32_t gtid) {
    Stack st;
    st.lds = lds + threadIdx.x;
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

__device__ __forceinline__ void flush_counts(const RayCounts& rc, unsigned long long* counters, uint32_t overflow,
                                             uint32_t* err) {
    const uint32_t c0 = wave_sum(rc.camera), c1 = wave_sum(rc.extension), c2 = wave_sum(rc.shadow),
                   c3 = wave_sum(rc.specular), c4 = wave_sum(rc.ts.pairs), c5 = wave_sum(rc.ts.prims),
                   of = wave_sum(overflow);
    if ((threadIdx.x & 63u) == 0u) {
        if (c0) atomicAdd(counters + 0, static_cast<unsigned long long>(c0));
        if (c1) atomicAdd(counters + 1, static_cast<unsigned long long>(c1));
        if (c2) atomicAdd(counters + 2, static_cast<unsigned long long>(c2));
        if (c3) atomicAdd(counters + 3, static_cast<unsigned long long>(c3));
        if (c4) atomicAdd(counters + 4, static_cast<unsigned long long>(c4));
        if (c5) atomicAdd(counters + 5, static_cast<unsigned long long>(c5));
        if (of) atomicOr(err, 1u);
    }
}

// ------------------------------------------------------------- trace --
template <bool GENERIC, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_trace_closest(SceneDev s, const mrt_ray* __restrict__ rays, uint32_t n,
                                                          mrt_hit* __restrict__ hits, StackCfg sc,
                                                          unsigned long long* counters, uint32_t* err) {
    __shared__ uint32_t lds[kStackLds * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    Stack st = make_stack(lds, sc, gtid);
    RayCounts rc{};
    for (uint32_t i = gtid; i < n; i += gridDim.x * kBlock) {
        const float4* rp = reinterpret_cast<const float4*>(rays + i);
        const float4 a = rp[0], b = rp[1];
        const RayT r = make_ray(f3{a.x, a.y, a.z}, f3{a.w, b.x, b.y}, __float_as_uint(b.w));
        const HitT h = closest_hit<GENERIC, COUNT>(s, r, b.z, st, rc.ts);
        reinterpret_cast<float4*>(hits)[i] =
            make_float4(h.id == MRT_ID_NONE ? 0.0f : h.t, __uint_as_float(h.id), h.u, h.v);
    }
    if (COUNT) flush_counts(rc, counters, st.overflow, err);
    else if (st.overflow) atomicOr(err, 1u);
}

template <bool GENERIC>
__global__ __launch_bounds__(kBlock) void k_trace_any(SceneDev s, const mrt_ray* __restrict__ rays, uint32_t n,
                                                      uint8_t* __restrict__ occ, StackCfg sc, uint32_t* err) {
    __shared__ uint32_t lds[kStackLds * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    Stack st = make_stack(lds, sc, gtid);
    TraceStats ts{};
    for (uint32_t i = gtid; i < n; i += gridDim.x * kBlock) {
        const float4* rp = reinterpret_cast<const float4*>(rays + i);
        const float4 a = rp[0], b = rp[1];
        const RayT r = make_ray(f3{a.x, a.y, a.z}, f3{a.w, b.x, b.y}, __float_as_uint(b.w));
        occ[i] = any_hit<GENERIC, false>(s, r, b.z, st, ts) ? 1 : 0;
    }
    if (st.overflow) atomicOr(err, 1u);
}

// ------------------------------------------------------------- camera --
__global__ __launch_bounds__(kBlock) void k_camera_rays(mrt_camera cam, const float* __restrict__ uvdd, uint32_t n,
                                                        mrt_ray* __restrict__ out) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 q = reinterpret_cast<const float4*>(uvdd)[i];
    f3 o, d;
    camera_ray(cam, q.x, q.y, q.z, q.w, &o, &d);
    float4* op = reinterpret_cast<float4*>(out + i);
    op[0] = make_float4(o.x, o.y, o.z, d.x);
    op[1] = make_float4(d.y, d.z, kRayLengthMax, __uint_as_float(MRT_ID_NONE));
}

// ------------------------------------------------------------- pairs --
// Ray i against primitive i of one category (known-answer tests).
__global__ __launch_bounds__(kBlock) void k_intersect_pairs(int cat, const void* __restrict__ prims,
                                                            const mrt_ray* __restrict__ rays, uint32_t n,
                                                            mrt_hit* __restrict__ closest, uint8_t* __restrict__ any) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const mrt_ray ry = rays[i];
    const RayT r = make_ray(ld3(ry.org), ld3(ry.dir), ry.exclude);
    const uint32_t id = MRT_MAKE_ID(cat, 0u);
    float t = 0.0f, u = 0.0f, v = 0.0f;
    bool hit = false, ah = false;
    if (cat == MRT_CAT_TRIANGLE || cat == MRT_CAT_LIGHT) {
        const mrt_triangle tr = static_cast<const mrt_triangle*>(prims)[i];
        TriG g;
        g.a = ld3(tr.a);
        g.ab = ld3(tr.ab);
        g.ac = ld3(tr.ac);
        if (id != r.excl) {
            hit = tri_test(g, r, &t, &u, &v) && t < ry.tmax;
            ah = tri_any(g, r, ry.tmax);
        }
    } else if (cat == MRT_CAT_SPHERE) {
        const mrt_sphere sp = static_cast<const mrt_sphere*>(prims)[i];
        if (sph_root(ld3(sp.center), sp.sq_radius, r, &t)) {
            hit = !(t < 1.0e-05f) && t < ry.tmax;
            ah = t > 1.0e-05f && t < ry.tmax;
        }
    } else {
        const mrt_plane pl = static_cast<const mrt_plane*>(prims)[i];
        if (id != r.excl && pla_t(ld3(pl.normal), ld3(pl.point), r, &t)) {
            hit = !(t < kEpsilon) && t < ry.tmax;
            ah = t > kEpsilon && t < ry.tmax;
        }
    }
    if (closest) {
        mrt_hit hh;
        hh.t = hit ? t : 0.0f;
        hh.id = hit ? id : MRT_ID_NONE;
        hh.u = hit ? u : 0.0f;
        hh.v = hit ? v : 0.0f;
        closest[i] = hh;
    }
    if (any) any[i] = ah ? 1 : 0;
}

// ------------------------------------------------------------- shade --
template <bool GENERIC, bool BRANCHES>
__global__ __launch_bounds__(kBlock) void k_shade_rays(SceneDev s, RenderParams p, const mrt_ray* __restrict__ rays,
                                                       const int32_t* __restrict__ depths, uint32_t n,
                                                       uint32_t sample, float* __restrict__ rgb,
                                                       uint8_t* __restrict__ flag, StackCfg sc, uint32_t* err) {
    __shared__ uint32_t lds[kStackLds * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    Stack st = make_stack(lds, sc, gtid);
    RayCounts rc{};
    uint32_t pendOverflow = 0;
    for (uint32_t i = gtid; i < n; i += gridDim.x * kBlock) {
        const mrt_ray ry = rays[i];
        bool hl = false;
        const f3 c = trace_path<GENERIC, BRANCHES, false>(s, p, ld3(ry.org), ld3(ry.dir), ry.exclude,
                                                          depths ? depths[i] : 1, i, sample, &hl, st, rc, 0,
                                                          &pendOverflow);
        rgb[3u * i + 0u] = c.x;
        rgb[3u * i + 1u] = c.y;
        rgb[3u * i + 2u] = c.z;
        flag[i] = hl ? 1 : 0;
    }
    if (st.overflow || pendOverflow) atomicOr(err, st.overflow ? 1u : 2u);
}

// ------------------------------------------------------------- render --
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

// One sample pass of Renderer::renderScene (Renderer.cpp:67-116) over the
// covered region: pixel sampler draws, Camera::generateRay, Shader::rayTrace,
// incrementalAvg into the bitmap and the FP32 radiance into accum.
template <bool GENERIC, bool BRANCHES, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_render_pass(SceneDev s, RenderParams p, uint32_t sample,
                                                        uint32_t* __restrict__ bitmap, float* __restrict__ accum,
                                                        StackCfg sc, unsigned long long* counters, uint32_t* err) {
    __shared__ uint32_t lds[kStackLds * kBlock];
    const uint32_t logical = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = logical * (kBlock / 64u) + (threadIdx.x >> 6);
    const uint32_t covW = 16u * p.bsx, covH = 16u * p.bsy;
    const uint32_t tilesX = (covW + 7u) >> 3;
    const uint32_t x = (wave % tilesX) * 8u + (lane & 7u);
    const uint32_t y = (wave / tilesX) * 8u + (lane >> 3);
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    Stack st = make_stack(lds, sc, gtid);
    RayCounts rc{};
    uint32_t pendOverflow = 0;
    if (x < covW && y < covH) {
        const uint32_t pid = y * p.width + x;
        float r1, r2;
        if (p.compat) {
            // HaltonSeq/Stratified::getSample(0) with the frame-wide counter in
            // single-thread claim order (HaltonSeq.cpp:14-22, Stratified.cpp:14-23).
            const uint32_t tx = x / p.bsx, ty = y / p.bsy;
            const uint32_t rank = p.tile_rank[ty * 16u + tx];
            const uint32_t ord = sample * p.pixels_per_pass + rank * p.bsx * p.bsy + (y - ty * p.bsy) * p.bsx +
                                 (x - tx * p.bsx);
            const uint32_t c0 = 2u * ord, c1 = c0 + 1u;
            const uint32_t D = p.s_pixel.domain_size;
            const bool halton_kind = p.s_pixel.kind == MRT_SAMPLER_HALTON_COMPAT;
            r1 = c0 < D ? (halton_kind ? halton(c0, 2u) : static_cast<float>(c0) / static_cast<float>(D)) : 1.0f;
            r2 = c1 < D ? (halton_kind ? halton(c1, 2u) : static_cast<float>(c1) / static_cast<float>(D)) : 1.0f;
        } else {
            r1 = draw(p.s_pixel, pid, sample, draw_dim(kSlotJitterU, 0u, 1u, 0u), p.spp);
            r2 = draw(p.s_pixel, pid, sample, draw_dim(kSlotJitterV, 0u, 1u, 0u), p.spp);
        }
        const float u = static_cast<float>(x) * p.inv_w;
        const float v = static_cast<float>(y) * p.inv_h;
        const float du = ((r1 - 0.5f) * 2.0f) * p.pix_w;
        const float dv = ((r2 - 0.5f) * 2.0f) * p.pix_h;
        f3 o, d;
        camera_ray(p.cam, u, v, du, dv, &o, &d);
        bool hl = false;
        const f3 c = trace_path<GENERIC, BRANCHES, COUNT>(s, p, o, d, MRT_ID_NONE, 1, pid, sample, &hl, st, rc, 0,
                                                          &pendOverflow);
        const uint32_t bi = y * p.real_width + x;
        bitmap[bi] = incremental_avg(c, bitmap[bi], sample + 1u);
        if (accum) {
            accum[3u * pid + 0u] += c.x;
            accum[3u * pid + 1u] += c.y;
            accum[3u * pid + 2u] += c.z;
        }
    }
    flush_counts(rc, counters, st.overflow | (pendOverflow << 1), err);
}

// ------------------------------------------------------------- launchers --
static uint32_t grid_for(uint32_t n, uint32_t cap_blocks) {
    uint32_t b = (n + kBlock - 1) / kBlock;
    if (cap_blocks && b > cap_blocks) b = cap_blocks;
    return b == 0 ? 1 : b;
}

hipError_t launch_trace_closest(const SceneDev& s, bool generic, const mrt_ray* rays, uint32_t n, mrt_hit* hits,
                                const StackCfg& sc, uint32_t cap_blocks, bool count, unsigned long long* counters,
                                uint32_t* err, hipStream_t stream) {
    const uint32_t g = grid_for(n, cap_blocks);
    if (generic) {
        if (count) k_trace_closest<true, true><<<g, kBlock, 0, stream>>>(s, rays, n, hits, sc, counters, err);
        else k_trace_closest<true, false><<<g, kBlock, 0, stream>>>(s, rays, n, hits, sc, counters, err);
    } else {
        if (count) k_trace_closest<false, true><<<g, kBlock, 0, stream>>>(s, rays, n, hits, sc, counters, err);
        else k_trace_closest<false, false><<<g, kBlock, 0, stream>>>(s, rays, n, hits, sc, counters, err);
    }
    return hipGetLastError();
}

hipError_t launch_trace_any(const SceneDev& s, bool generic, const mrt_ray* rays, uint32_t n, uint8_t* occ,
                            const StackCfg& sc, uint32_t cap_blocks, uint32_t* err, hipStream_t stream) {
    const uint32_t g = grid_for(n, cap_blocks);
    if (generic) k_trace_any<true><<<g, kBlock, 0, stream>>>(s, rays, n, occ, sc, err);
    else k_trace_any<false><<<g, kBlock, 0, stream>>>(s, rays, n, occ, sc, err);
    return hipGetLastError();
}

hipError_t launch_camera_rays(const mrt_camera& cam, const float* uvdd, uint32_t n, mrt_ray* out,
                              hipStream_t stream) {
    k_camera_rays<<<grid_for(n, 0), kBlock, 0, stream>>>(cam, uvdd, n, out);
    return hipGetLastError();
}

hipError_t launch_intersect_pairs(int cat, const void* prims, const mrt_ray* rays, uint32_t n, mrt_hit* closest,
                                  uint8_t* any, hipStream_t stream) {
    k_intersect_pairs<<<grid_for(n, 0), kBlock, 0, stream>>>(cat, prims, rays, n, closest, any);
    return hipGetLastError();
}

hipError_t launch_shade_rays(const SceneDev& s, bool generic, bool branches, const RenderParams& p,
                             const mrt_ray* rays, const int32_t* depths, uint32_t n, uint32_t sample, float* rgb,
                             uint8_t* flag, const StackCfg& sc, uint32_t cap_blocks, uint32_t* err,
                             hipStream_t stream) {
    const uint32_t g = grid_for(n, cap_blocks);
    if (generic) {
        if (branches) k_shade_rays<true, true><<<g, kBlock, 0, stream>>>(s, p, rays, depths, n, sample, rgb, flag, sc, err);
        else k_shade_rays<true, false><<<g, kBlock, 0, stream>>>(s, p, rays, depths, n, sample, rgb, flag, sc, err);
    } else {
        if (branches) k_shade_rays<false, true><<<g, kBlock, 0, stream>>>(s, p, rays, depths, n, sample, rgb, flag, sc, err);
        else k_shade_rays<false, false><<<g, kBlock, 0, stream>>>(s, p, rays, depths, n, sample, rgb, flag, sc, err);
    }
    return hipGetLastError();
}

uint32_t render_pass_blocks(const RenderParams& p) {
    const uint32_t covW = 16u * p.bsx, covH = 16u * p.bsy;
    const uint32_t tiles = ((covW + 7u) / 8u) * ((covH + 7u) / 8u);
    return (tiles + 3u) / 4u;
}

hipError_t launch_render_pass(const SceneDev& s, bool generic, bool branches, bool count, const RenderParams& p,
                              uint32_t sample, uint32_t* bitmap, float* accum, const StackCfg& sc,
                              unsigned long long* counters, uint32_t* err, hipStream_t stream) {
    const uint32_t g = render_pass_blocks(p);
#define MRT_RENDER(G, B, C) \
    k_render_pass<G, B, C><<<g, kBlock, 0, stream>>>(s, p, sample, bitmap, accum, sc, counters, err)
    if (generic) {
        if (branches) { if (count) MRT_RENDER(true, true, true); else MRT_RENDER(true, true, false); }
        else { if (count) MRT_RENDER(true, false, true); else MRT_RENDER(true, false, false); }
    } else {
        if (branches) { if (count) MRT_RENDER(false, true, true); else MRT_RENDER(false, true, false); }
        else { if (count) MRT_RENDER(false, false, true); else MRT_RENDER(false, false, false); }
    }
#undef MRT_RENDER
    return hipGetLastError();
}

}  // namespace mrt
