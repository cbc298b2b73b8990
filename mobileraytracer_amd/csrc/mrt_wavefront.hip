// Wavefront path tracer: one sample pass runs as
//   generate -> (extend -> shade -> shadow) x RayDepthMax -> finish
// kernels joined by queues in HBM.  shade() appends extension and shadow rays
// to compacted queues (one atomic per wave: ballot + popcount + mbcnt), so
// every traversal launch runs full waves of live rays, and each kernel keeps
// its own register budget instead of the megakernel's union of all of them.
//
// Per-path state (48 bytes per covered pixel):
//   sA = (T.xyz, flags)   flags: bit 0 found, bit 1 ends, bits 8..15 depth
//   sB = (acc_all.xyz, pixel id)
//   sC = (acc_prefix.xyz, 0)
// with the chain semantics of mrt_path.hpp (PathTracer.cpp:25-164): the
// shadow kernel adds T_d * Ld_d and records the first vertex with
// any(Ld > 0); finish() picks acc_prefix when the chain ended on an emitter.
// Valid for the PathTracer with one light sample per vertex and no
// reflection/transmission materials (the host checks; other cases use the
// megakernel).
#include <hipcub/hipcub.hpp>

#include "mrt_device_util.hpp"
#include "mrt_wavefront.hpp"

namespace mrt {

constexpr uint32_t kFound = 1u, kEnds = 2u;

__device__ __forceinline__ uint32_t flags_depth(uint32_t f) { return (f >> 8) & 0xFFu; }
__device__ __forceinline__ uint32_t with_depth(uint32_t f, uint32_t d) { return (f & ~0xFF00u) | (d << 8); }

__device__ __forceinline__ uint32_t spread5(uint32_t v) {
    v &= 31u;
    v = (v | (v << 8)) & 0x0300Fu;
    v = (v | (v << 4)) & 0x030C3u;
    v = (v | (v << 2)) & 0x09249u;
    return v;
}

// Sort key of an extension ray: Morton cell (5 bits per axis) of the origin
// and the direction octant.
__device__ __forceinline__ uint32_t ray_key(const WfBuffers& b, f3 o, f3 d) {
    const float fx = (o.x - b.box_mn[0]) * b.box_inv[0];
    const float fy = (o.y - b.box_mn[1]) * b.box_inv[1];
    const float fz = (o.z - b.box_mn[2]) * b.box_inv[2];
    const uint32_t x = static_cast<uint32_t>(fminf(fmaxf(fx, 0.0f), 31.0f));
    const uint32_t y = static_cast<uint32_t>(fminf(fmaxf(fy, 0.0f), 31.0f));
    const uint32_t z = static_cast<uint32_t>(fminf(fmaxf(fz, 0.0f), 31.0f));
    const uint32_t m = (spread5(x) << 2) | (spread5(y) << 1) | spread5(z);
    const uint32_t oct = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
    return (m << 3) | oct;
}

// Entries [begin, end) of a queue of n entries owned by this workgroup: the
// workgroups an XCD receives (every 8th) own one contiguous range.
__device__ __forceinline__ void block_range(uint32_t n, uint32_t* begin, uint32_t* end) {
    const uint32_t nb = gridDim.x;
    const uint32_t lb = xcd_remap(blockIdx.x, nb);
    const uint32_t chunk = (((n + nb - 1u) / nb) + 63u) & ~63u;
    *begin = min(n, lb * chunk);
    *end = min(n, *begin + chunk);
}

// Reserves `want ? 1 : 0` slots of a queue per lane with one atomic per wave;
// returns the lane's slot (valid only where want is true).
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool want) {
    const unsigned long long mask = __ballot(want);
    const uint32_t total = static_cast<uint32_t>(__popcll(mask));
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t rank = static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
    const uint32_t leader = static_cast<uint32_t>(__ffsll(static_cast<long long>(mask))) - 1u;
    uint32_t base = 0;
    if (total && lane == leader) base = atomicAdd(counter, total);
    base = __shfl(base, static_cast<int>(leader), 64);
    return base + rank;
}

// ------------------------------------------------------------- generate --
__global__ __launch_bounds__(kBlock) void k_wf_generate(RenderParams p, uint32_t sample, uint32_t slots,
                                                        WfBuffers b) {
    const uint32_t logical = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t slot = logical * kBlock + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        b.counts[0] = slots;
        b.counts[3] = 0u;
        b.counts[4] = 0u;
    }
    if (slot >= slots) return;
    uint32_t x, y;
    slot_pixel(p, slot, &x, &y);
    float4* q = b.queue_a + 2u * slot;
    if (x >= 16u * p.bsx || y >= 16u * p.bsy) {
        q[1] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0xFFFFFFFFu));
        return;
    }
    const uint32_t pid = y * p.width + x;
    f3 o, d;
    primary_ray(p, x, y, pid, sample, &o, &d);
    q[0] = make_float4(o.x, o.y, o.z, d.x);
    q[1] = make_float4(d.y, d.z, __uint_as_float(MRT_ID_NONE), __uint_as_float(slot));
    b.sA[slot] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(with_depth(0u, 1u)));
    b.sB[slot] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(pid));
    b.sC[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ------------------------------------------------------------- extend --
template <bool GENERIC, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_wf_extend(SceneDev s, const float4* __restrict__ queue,
                                                      const uint32_t* __restrict__ count, float4* __restrict__ hits,
                                                      uint32_t* clear_a, uint32_t* clear_b, StackCfg sc,
                                                      unsigned long long* counters, uint32_t* err) {
    __shared__ uint32_t lds[kStackLds * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    // The queues shade() fills next were drained by the previous launches.
    if (gtid == 0) {
        *clear_a = 0u;
        *clear_b = 0u;
    }
    Stack st = make_stack(lds, sc, gtid);
    RayCounts rc{};
    uint32_t begin, end;
    block_range(*count, &begin, &end);
    for (uint32_t i = begin + threadIdx.x; i < end; i += kBlock) {
        const float4 a = queue[2u * i], c = queue[2u * i + 1u];
        if (__float_as_uint(c.w) == 0xFFFFFFFFu) continue;
        const RayT r = make_ray(f3{a.x, a.y, a.z}, f3{a.w, c.x, c.y}, __float_as_uint(c.z));
        const HitT h = closest_hit<GENERIC, COUNT>(s, r, kRayLengthMax, st, rc.ts);
        hits[i] = make_float4(h.t, __uint_as_float(h.id), h.u, h.v);
    }
    if (COUNT) flush_counts<true>(rc, counters, st.overflow, err);
    else if (st.overflow) atomicOr(err, 1u);
}

// Persistent closest-hit over the queue for BVH scenes of triangles and
// lights: a lane takes the next queue entry as soon as its ray is done
// (refills are batched per wave: at least 8 idle lanes, or all of them), so
// the wave never waits for its slowest ray.
template <bool COUNT>
__global__ __launch_bounds__(kBlock) void k_wf_extend_p(SceneDev s, const float4* __restrict__ queue,
                                                        const uint32_t* __restrict__ count, float4* __restrict__ hits,
                                                        uint32_t* clear_a, uint32_t* clear_b, uint32_t* clear_c,
                                                        uint32_t* fetch, StackCfg sc, unsigned long long* counters,
                                                        uint32_t* err) {
    __shared__ uint32_t lds[kStackLds * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    if (gtid == 0) {
        *clear_a = 0u;
        *clear_b = 0u;
        *clear_c = 0u;
    }
    Stack st = make_stack(lds, sc, gtid);
    RayCounts rc{};
    const uint32_t n = *count;
    Trav t;
    uint32_t idx = 0;
    bool has = false, exhausted = false;
    while (true) {
        const bool need = !has && !exhausted;
        const unsigned long long needMask = __ballot(need);
        if (needMask != 0ull && (__popcll(needMask) >= 8 || __ballot(has) == 0ull)) {
            const uint32_t slot = wave_append(fetch, need);
            if (need) {
                if (slot < n) {
                    const float4 a = queue[2u * slot], c = queue[2u * slot + 1u];
                    if (__float_as_uint(c.w) != 0xFFFFFFFFu) {
                        idx = slot;
                        t.r = make_ray(f3{a.x, a.y, a.z}, f3{a.w, c.x, c.y}, __float_as_uint(c.z));
                        trav_begin(s, t, st, kRayLengthMax);
                        has = true;
                    }
                } else {
                    exhausted = true;
                }
            }
        }
        if (__ballot(has || !exhausted) == 0ull) break;
        if (has && step_closest<COUNT>(s, t, st, rc.ts)) {
            hits[idx] = make_float4(t.h.t, __uint_as_float(t.h.id), t.h.u, t.h.v);
            has = false;
        }
    }
    if (COUNT) flush_counts<true>(rc, counters, st.overflow, err);
    else if (st.overflow) atomicOr(err, 1u);
}

// ------------------------------------------------------------- shade --
// One PathTracer vertex per queue entry (PathTracer.cpp:25-133).
__global__ __launch_bounds__(kBlock) void k_wf_shade(SceneDev s, RenderParams p, uint32_t sample,
                                                     const float4* __restrict__ queue_in,
                                                     const uint32_t* __restrict__ count_in,
                                                     const float4* __restrict__ hits, WfBuffers b,
                                                     float4* __restrict__ queue_out, uint32_t* count_out,
                                                     uint32_t* count_shadow, uint32_t* clear_fetch,
                                                     unsigned long long* counters) {
    RayCounts rc{};
    if (blockIdx.x == 0 && threadIdx.x == 0) *clear_fetch = 0u;
    const uint32_t n = *count_in;
    const uint32_t stride = gridDim.x * kBlock;
    // Every lane of a wave runs the same number of iterations so that the
    // wave-wide appends below see all lanes.
    const uint32_t rounds = (n + stride - 1u) / stride;
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    for (uint32_t k = 0; k < rounds; ++k) {
        const uint32_t i = gtid + k * stride;
        bool wantExt = false, wantShadow = false;
        f3 nextO{}, nextD{}, shD{}, Ld{}, Tv{};
        uint32_t excl = MRT_ID_NONE, slot = 0xFFFFFFFFu;
        float dist = 0.0f;
        if (i < n) {
            const float4 a = queue_in[2u * i], c = queue_in[2u * i + 1u];
            slot = __float_as_uint(c.w);
            if (slot != 0xFFFFFFFFu) {
                const RayT r = make_ray(f3{a.x, a.y, a.z}, f3{a.w, c.x, c.y}, __float_as_uint(c.z));
                const float4 hv = hits[i];
                HitT h;
                h.t = hv.x;
                h.id = __float_as_uint(hv.y);
                h.u = hv.z;
                h.v = hv.w;
                float4 sa = b.sA[slot];
                uint32_t flags = __float_as_uint(sa.w);
                const uint32_t depth = flags_depth(flags);
                if (depth == 1u) ++rc.camera;
                else ++rc.extension;
                if (h.id != MRT_ID_NONE) {
                    const f3 T = f3{sa.x, sa.y, sa.z};
                    const Surf sf = surface(s, r, h);
                    const Shading m = shading_of(s, sf);
                    const DrawCtx dc{__float_as_uint(b.sB[slot].w), sample, 0u};
                    if (m.emitter) {
                        float4 sbv = b.sB[slot];
                        const f3 acc = add(f3{sbv.x, sbv.y, sbv.z}, mul(T, m.le));
                        b.sB[slot] = make_float4(acc.x, acc.y, acc.z, sbv.w);
                        flags |= kEnds;
                        b.sA[slot] = make_float4(sa.x, sa.y, sa.z, __uint_as_float(flags));
                    } else if (any_gt0(m.kd)) {
                        if (s.n_lights > 0u) {
                            f3 lc;
                            if (light_sample(s, p, sf, static_cast<int32_t>(depth), dc, 0u, &shD, &dist, &lc)) {
                                wantShadow = true;
                                Ld = divs(mul(lc, m.kd), static_cast<float>(p.samples_light));
                                Tv = T;
                                excl = sf.excl;
                                nextO = sf.point;
                            }
                        }
                        bool cont = static_cast<int32_t>(depth) <= p.depth_min;
                        if (!cont) cont = draw_at(p, p.s_rr, dc, kSlotRR, 0, static_cast<int32_t>(depth)) > 0.5f;
                        if (cont && static_cast<int32_t>(depth) + 1 <= p.depth_max) {
                            nextD = cosine_hemisphere(p, sf.normal, dc, static_cast<int32_t>(depth));
                            f3 Tn = mul(T, m.kd);
                            if (static_cast<int32_t>(depth) > p.depth_min) Tn = divs(Tn, 0.5f * 0.5f);
                            wantExt = true;
                            excl = sf.excl;
                            nextO = sf.point;
                            b.sA[slot] = make_float4(Tn.x, Tn.y, Tn.z, __uint_as_float(with_depth(flags, depth + 1u)));
                        }
                    }
                }
            }
        }
        const uint32_t qe = wave_append(count_out, wantExt);
        if (wantExt) {
            queue_out[2u * qe] = make_float4(nextO.x, nextO.y, nextO.z, nextD.x);
            queue_out[2u * qe + 1u] = make_float4(nextD.y, nextD.z, __uint_as_float(excl), __uint_as_float(slot));
            if (b.sort) {
                const uint32_t key = ray_key(b, nextO, nextD);
                b.keys[qe] = key;
                atomicAdd(b.hist + key, 1u);
            }
        }
        const uint32_t qs = wave_append(count_shadow, wantShadow);
        if (wantShadow) {
            float4* e = b.shadow + 4u * qs;
            e[0] = make_float4(nextO.x, nextO.y, nextO.z, shD.x);
            e[1] = make_float4(shD.y, shD.z, dist, __uint_as_float(excl));
            e[2] = make_float4(Ld.x, Ld.y, Ld.z, __uint_as_float(slot));
            e[3] = make_float4(Tv.x, Tv.y, Tv.z, 0.0f);
        }
    }
    flush_counts<false>(rc, counters, 0u, nullptr);
}

// ------------------------------------------------------------- shadow --
// Shader::shadowTrace of each queued light sample; a visible sample adds
// T_d * Ld_d to acc_all and may mark the first vertex with any(Ld > 0).
template <bool GENERIC, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_wf_shadow(SceneDev s, const uint32_t* __restrict__ count, WfBuffers b,
                                                      StackCfg sc, unsigned long long* counters, uint32_t* err) {
    __shared__ uint32_t lds[kStackLds * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    Stack st = make_stack(lds, sc, gtid);
    RayCounts rc{};
    uint32_t begin, end;
    block_range(*count, &begin, &end);
    for (uint32_t i = begin + threadIdx.x; i < end; i += kBlock) {
        const float4* e = b.shadow + 4u * i;
        const float4 a = e[0], c = e[1];
        ++rc.shadow;
        const RayT r = make_ray(f3{a.x, a.y, a.z}, f3{a.w, c.x, c.y}, __float_as_uint(c.w));
        if (any_hit<GENERIC, COUNT>(s, r, c.z, st, rc.ts)) continue;
        const float4 l = e[2], tv = e[3];
        const uint32_t slot = __float_as_uint(l.w);
        const f3 Ld = f3{l.x, l.y, l.z};
        const float4 sbv = b.sB[slot];
        const f3 acc = add(f3{sbv.x, sbv.y, sbv.z}, mul(f3{tv.x, tv.y, tv.z}, Ld));
        b.sB[slot] = make_float4(acc.x, acc.y, acc.z, sbv.w);
        const float4 sa = b.sA[slot];
        const uint32_t flags = __float_as_uint(sa.w);
        if (!(flags & kFound) && any_gt0(Ld)) {
            b.sA[slot] = make_float4(sa.x, sa.y, sa.z, __uint_as_float(flags | kFound));
            b.sC[slot] = make_float4(acc.x, acc.y, acc.z, 0.0f);
        }
    }
    if (COUNT) flush_counts<true>(rc, counters, st.overflow, err);
    else flush_counts<false>(rc, counters, st.overflow, err);
}

// Applies a visible light sample to its path (see k_wf_shadow).
__device__ __forceinline__ void apply_light(const WfBuffers& b, const float4* e) {
    const float4 l = e[2], tv = e[3];
    const uint32_t slot = __float_as_uint(l.w);
    const f3 Ld = f3{l.x, l.y, l.z};
    const float4 sbv = b.sB[slot];
    const f3 acc = add(f3{sbv.x, sbv.y, sbv.z}, mul(f3{tv.x, tv.y, tv.z}, Ld));
    b.sB[slot] = make_float4(acc.x, acc.y, acc.z, sbv.w);
    const float4 sa = b.sA[slot];
    const uint32_t flags = __float_as_uint(sa.w);
    if (!(flags & kFound) && any_gt0(Ld)) {
        b.sA[slot] = make_float4(sa.x, sa.y, sa.z, __uint_as_float(flags | kFound));
        b.sC[slot] = make_float4(acc.x, acc.y, acc.z, 0.0f);
    }
}

// Persistent shadow rays for BVH scenes of triangles and lights (refill as
// in k_wf_extend_p).
template <bool COUNT>
__global__ __launch_bounds__(kBlock) void k_wf_shadow_p(SceneDev s, const uint32_t* __restrict__ count, WfBuffers b,
                                                        uint32_t* fetch, StackCfg sc, unsigned long long* counters,
                                                        uint32_t* err) {
    __shared__ uint32_t lds[kStackLds * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    Stack st = make_stack(lds, sc, gtid);
    RayCounts rc{};
    const uint32_t n = *count;
    Trav t;
    uint32_t idx = 0;
    float dist = 0.0f;
    bool has = false, exhausted = false;
    while (true) {
        const bool need = !has && !exhausted;
        const unsigned long long needMask = __ballot(need);
        if (needMask != 0ull && (__popcll(needMask) >= 8 || __ballot(has) == 0ull)) {
            const uint32_t slot = wave_append(fetch, need);
            if (need) {
                if (slot < n) {
                    const float4* e = b.shadow + 4u * slot;
                    const float4 a = e[0], c = e[1];
                    idx = slot;
                    t.r = make_ray(f3{a.x, a.y, a.z}, f3{a.w, c.x, c.y}, __float_as_uint(c.w));
                    dist = c.z;
                    t.node = kRootPending;
                    st.sp = 0;
                    ++rc.shadow;
                    has = true;
                } else {
                    exhausted = true;
                }
            }
        }
        if (__ballot(has || !exhausted) == 0ull) break;
        if (has) {
            bool occluded = false;
            if (step_any<COUNT>(s, t, dist, st, rc.ts, &occluded)) {
                if (!occluded) apply_light(b, b.shadow + 4u * idx);
                has = false;
            }
        }
    }
    if (COUNT) flush_counts<true>(rc, counters, st.overflow, err);
    else flush_counts<false>(rc, counters, st.overflow, err);
}

// ------------------------------------------------------------- reorder --
// Counting-sort scatter of shade()'s output into the traced queue, after an
// exclusive scan turned the histogram into offsets; clears the histogram for
// the next bounce.  The order inside a bin varies from run to run; no path's
// result depends on its queue position.
__global__ __launch_bounds__(kBlock) void k_wf_scatter(const uint32_t* __restrict__ count, WfBuffers b) {
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t n = *count;
    for (uint32_t i = gtid; i < n; i += stride) {
        const uint32_t pos = atomicAdd(b.offs + b.keys[i], 1u);
        b.queue_a[2u * pos] = b.queue_b[2u * i];
        b.queue_a[2u * pos + 1u] = b.queue_b[2u * i + 1u];
    }
    for (uint32_t k = gtid; k < kSortBins; k += stride) b.hist[k] = 0u;
}

// ------------------------------------------------------------- finish --
__global__ __launch_bounds__(kBlock) void k_wf_finish(RenderParams p, uint32_t sample, uint32_t slots, WfBuffers b,
                                                      uint32_t* __restrict__ bitmap, float* __restrict__ accum) {
    const uint32_t slot = blockIdx.x * kBlock + threadIdx.x;
    if (slot >= slots) return;
    uint32_t x, y;
    slot_pixel(p, slot, &x, &y);
    if (x >= 16u * p.bsx || y >= 16u * p.bsy) return;
    const float4 sa = b.sA[slot], sbv = b.sB[slot];
    const uint32_t flags = __float_as_uint(sa.w);
    f3 c = f3{sbv.x, sbv.y, sbv.z};
    if ((flags & kEnds) && (flags & kFound)) {
        const float4 scv = b.sC[slot];
        c = f3{scv.x, scv.y, scv.z};
    }
    const uint32_t pid = __float_as_uint(sbv.w);
    const uint32_t bi = y * p.real_width + x;
    bitmap[bi] = incremental_avg(c, bitmap[bi], sample + 1u);
    if (accum) {
        accum[3u * pid + 0u] += c.x;
        accum[3u * pid + 1u] += c.y;
        accum[3u * pid + 2u] += c.z;
    }
}

// ------------------------------------------------------------- launcher --
size_t wavefront_scan_bytes() {
    size_t bytes = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, static_cast<const uint32_t*>(nullptr),
                                           static_cast<uint32_t*>(nullptr), kSortBins);
    return bytes;
}

uint32_t wavefront_slots(const RenderParams& p) {
    const uint32_t covW = 16u * p.bsx, covH = 16u * p.bsy;
    return ((covW + 7u) / 8u) * ((covH + 7u) / 8u) * 64u;
}

hipError_t launch_wavefront_pass(const SceneDev& s, bool generic, bool persistent, bool count, const RenderParams& p,
                                 uint32_t sample, uint32_t* bitmap, float* accum, const WfBuffers& b,
                                 uint32_t trace_blocks, const StackCfg& sc, unsigned long long* counters,
                                 uint32_t* err, hipStream_t stream) {
    const uint32_t slots = wavefront_slots(p);
    const uint32_t gblocks = (slots + kBlock - 1u) / kBlock;
    // counts[0] = entries of queue_a (set by generate), counts[1] = of
    // queue_b, counts[2] = shadow entries; extend clears the two shade fills.
    // counts[3] / counts[4] are the fetch cursors of the persistent extend /
    // shadow kernels: generate clears both, shade() clears extend's for the
    // next bounce, extend clears shadow's.
    k_wf_generate<<<gblocks, kBlock, 0, stream>>>(p, sample, slots, b);
    float4* qin = b.queue_a;
    float4* qout = b.queue_b;
    uint32_t* cin = b.counts;
    uint32_t* cout = b.counts + 1;
    uint32_t* csh = b.counts + 2;
    uint32_t* fext = b.counts + 3;
    uint32_t* fsh = b.counts + 4;
    for (int32_t depth = 1; depth <= p.depth_max; ++depth) {
        if (persistent) {
            if (count) k_wf_extend_p<true><<<trace_blocks, kBlock, 0, stream>>>(s, qin, cin, b.hits, cout, csh, fsh, fext, sc, counters, err);
            else k_wf_extend_p<false><<<trace_blocks, kBlock, 0, stream>>>(s, qin, cin, b.hits, cout, csh, fsh, fext, sc, counters, err);
        } else {
#define MRT_EXTEND(G, C) \
    k_wf_extend<G, C><<<trace_blocks, kBlock, 0, stream>>>(s, qin, cin, b.hits, cout, csh, sc, counters, err)
            if (generic) {
                if (count) MRT_EXTEND(true, true); else MRT_EXTEND(true, false);
            } else {
                if (count) MRT_EXTEND(false, true); else MRT_EXTEND(false, false);
            }
#undef MRT_EXTEND
        }
        k_wf_shade<<<trace_blocks, kBlock, 0, stream>>>(s, p, sample, qin, cin, b.hits, b, qout, cout, csh, fext,
                                                        counters);
        if (persistent) {
            if (count) k_wf_shadow_p<true><<<trace_blocks, kBlock, 0, stream>>>(s, csh, b, fsh, sc, counters, err);
            else k_wf_shadow_p<false><<<trace_blocks, kBlock, 0, stream>>>(s, csh, b, fsh, sc, counters, err);
        } else if (generic) {
            if (count) k_wf_shadow<true, true><<<trace_blocks, kBlock, 0, stream>>>(s, csh, b, sc, counters, err);
            else k_wf_shadow<true, false><<<trace_blocks, kBlock, 0, stream>>>(s, csh, b, sc, counters, err);
        } else {
            if (count) k_wf_shadow<false, true><<<trace_blocks, kBlock, 0, stream>>>(s, csh, b, sc, counters, err);
            else k_wf_shadow<false, false><<<trace_blocks, kBlock, 0, stream>>>(s, csh, b, sc, counters, err);
        }
        if (b.sort && depth < p.depth_max) {
            size_t bytes = b.scan_bytes;
            hipError_t e = hipcub::DeviceScan::ExclusiveSum(b.scan_temp, bytes, b.hist, b.offs, kSortBins, stream);
            if (e != hipSuccess) return e;
            k_wf_scatter<<<trace_blocks, kBlock, 0, stream>>>(cout, b);
        } else {
            float4* tq = qin;
            qin = qout;
            qout = tq;
        }
        uint32_t* tc = cin;
        cin = cout;
        cout = tc;
    }
    k_wf_finish<<<gblocks, kBlock, 0, stream>>>(p, sample, slots, b, bitmap, accum);
    return hipGetLastError();
}

}  // namespace mrt
