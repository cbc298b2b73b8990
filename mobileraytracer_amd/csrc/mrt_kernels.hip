// gfx950 kernels of the ray-scene path (traversal entry points and the
// megakernel renderer) and their host launchers.
//
// Launch shape: 256-thread workgroups (4 wave64s).  Each lane traces its own
// ray; the per-lane traversal stack keeps its first kStackLds entries in LDS
// (16 KiB per workgroup) and spills deeper entries to a global area of
// cap_spill entries per lane, sized from the BVH depth by the host.
// Render passes map one wave to an 8x8 pixel block and renumber workgroups so
// that each of the 8 XCDs (which receive workgroups round-robin) works on a
// contiguous band of the image, keeping neighbouring rays in one L2.
#include <mutex>

#include "mrt_device_util.hpp"

namespace mrt {

// Scratch of mrt_trace_closest_noalloc (see noalloc_scratch).  Host code
// never uses this name as an address: it takes each GPU's copy from
// hipGetSymbolAddress.
static __device__ __align__(256) uint32_t scratch[kNoAllocScratchWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(int device, uint32_t** out) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    k_fill_u32<<<(n + kBlock - 1) / kBlock, kBlock, 0, stream>>>(p, n, v);
    return hipGetLastError();
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
    if (COUNT) flush_counts<true>(rc, counters, st.overflow, err);
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
        const int32_t depth = depths ? depths[i] : 1;
        const f3 c = simple_shader(p.shader)
                         ? trace_simple<GENERIC, false>(s, p, ld3(ry.org), ld3(ry.dir), ry.exclude, depth, i, sample,
                                                        &hl, st, rc, 0)
                         : trace_path<GENERIC, BRANCHES, false>(s, p, ld3(ry.org), ld3(ry.dir), ry.exclude, depth, i,
                                                                sample, &hl, st, rc, 0, &pendOverflow);
        rgb[3u * i + 0u] = c.x;
        rgb[3u * i + 1u] = c.y;
        rgb[3u * i + 2u] = c.z;
        flag[i] = hl ? 1 : 0;
    }
    if (st.overflow || pendOverflow) atomicOr(err, st.overflow ? 1u : 2u);
}

// ------------------------------------------------------------- render --
// One sample pass of Renderer::renderScene (Renderer.cpp:67-116) over the
// covered region: pixel sampler draws, Camera::generateRay, Shader::rayTrace,
// incrementalAvg into the bitmap and the FP32 radiance into accum.
template <bool GENERIC, bool BRANCHES, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_render_pass(SceneDev s, RenderParams p, uint32_t sample,
                                                        uint32_t* __restrict__ bitmap, float* __restrict__ accum,
                                                        StackCfg sc, unsigned long long* counters, uint32_t* err) {
    __shared__ uint32_t lds[kStackLds * kBlock];
    const uint32_t logical = xcd_remap(blockIdx.x, gridDim.x);
    uint32_t x, y;
    slot_pixel(p, logical * kBlock + threadIdx.x, &x, &y);
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    Stack st = make_stack(lds, sc, gtid);
    RayCounts rc{};
    uint32_t pendOverflow = 0;
    if (pixel_covered(p, x, y)) {
        const uint32_t pid = y * p.width + x;
        f3 o, d;
        primary_ray(p, x, y, pid, sample, &o, &d);
        bool hl = false;
        const f3 c = simple_shader(p.shader)
                         ? trace_simple<GENERIC, COUNT>(s, p, o, d, MRT_ID_NONE, 1, pid, sample, &hl, st, rc, 0)
                         : trace_path<GENERIC, BRANCHES, COUNT>(s, p, o, d, MRT_ID_NONE, 1, pid, sample, &hl, st, rc,
                                                                0, &pendOverflow);
        const uint32_t bi = y * p.real_width + x;
        bitmap[bi] = incremental_avg(c, bitmap[bi], sample + 1u);
        if (accum) {
            accum[3u * pid + 0u] += c.x;
            accum[3u * pid + 1u] += c.y;
            accum[3u * pid + 2u] += c.z;
        }
    }
    flush_counts<COUNT>(rc, counters, st.overflow | (pendOverflow << 1), err);
}

// ------------------------------------------------------------- AOVs --
// Auxiliary outputs of a frame for a denoiser (no counterpart in the
// reference): per covered pixel, over the camera rays of samples [first,
// spp) (the render's own jitter), the mean first-hit albedo (the material's
// kD, the vertex colour without a material, Le on an emitter), shading normal
// and hit distance; a miss contributes zeros.  aov: 7 floats per pixel of the
// width x height frame.
template <bool GENERIC>
__global__ __launch_bounds__(kBlock) void k_aov(SceneDev s, RenderParams p, uint32_t first, uint32_t spp,
                                                float* __restrict__ aov, StackCfg sc, uint32_t* err) {
    __shared__ uint32_t lds[kStackLds * kBlock];
    const uint32_t logical = xcd_remap(blockIdx.x, gridDim.x);
    uint32_t x, y;
    slot_pixel(p, logical * kBlock + threadIdx.x, &x, &y);
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    Stack st = make_stack(lds, sc, gtid);
    TraceStats ts{};
    if (pixel_covered(p, x, y)) {
        const uint32_t pid = y * p.width + x;
        f3 alb{}, nrm{};
        float dep = 0.0f;
        for (uint32_t sample = first; sample < spp; ++sample) {
            f3 o, d;
            primary_ray(p, x, y, pid, sample, &o, &d);
            const RayT r = make_ray(o, d, MRT_ID_NONE);
            const HitT h = closest_hit<GENERIC, false>(s, r, kRayLengthMax, st, ts);
            if (h.id == MRT_ID_NONE) continue;
            const Surf sf = surface(s, r, h);
            const Shading m = shading_of(s, sf);
            alb = add(alb, m.emitter ? m.le : m.kd);
            nrm = add(nrm, sf.normal);
            dep += h.t;
        }
        const float inv = 1.0f / static_cast<float>(spp - first);
        float* o7 = aov + 7ull * pid;
        o7[0] = alb.x * inv;
        o7[1] = alb.y * inv;
        o7[2] = alb.z * inv;
        o7[3] = nrm.x * inv;
        o7[4] = nrm.y * inv;
        o7[5] = nrm.z * inv;
        o7[6] = dep * inv;
    }
    if (st.overflow) atomicOr(err, 1u);
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

// ::MobileRT::intersect(AABB, Ray) (AABB.cpp:15-35) of box i and ray i.
__global__ __launch_bounds__(kBlock) void k_box_rays(const float* __restrict__ boxes6,
                                                     const mrt_ray* __restrict__ rays, uint32_t n,
                                                     uint8_t* __restrict__ hit, float* __restrict__ entry) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const mrt_ray ry = rays[i];
    const RayT r = make_ray(ld3(ry.org), ld3(ry.dir), ry.exclude);
    const float* b = boxes6 + 6u * i;
    float e = 0.0f;
    const bool h = slab(f3{b[0], b[1], b[2]}, f3{b[3], b[4], b[5]}, r.o, r.inv, &e);
    hit[i] = h ? 1u : 0u;
    if (entry) entry[i] = e;
}

hipError_t launch_box_rays(const float* boxes6, const mrt_ray* rays, uint32_t n, uint8_t* hit, float* entry,
                           hipStream_t stream) {
    k_box_rays<<<grid_for(n, 0), kBlock, 0, stream>>>(boxes6, rays, n, hit, entry);
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

hipError_t launch_aov(const SceneDev& s, bool generic, const RenderParams& p, uint32_t first, uint32_t spp, float* aov,
                      const StackCfg& sc, uint32_t* err, hipStream_t stream) {
    const uint32_t g = render_pass_blocks(p);
    if (generic) k_aov<true><<<g, kBlock, 0, stream>>>(s, p, first, spp, aov, sc, err);
    else k_aov<false><<<g, kBlock, 0, stream>>>(s, p, first, spp, aov, sc, err);
    return hipGetLastError();
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
