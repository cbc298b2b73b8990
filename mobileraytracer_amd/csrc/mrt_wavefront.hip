// Wavefront path tracer with path regeneration.  Every covered pixel owns
// one path slot and runs its samples one after the other; the frame is a
// loop of iterations
//   trace -> regen -> [reorder] -> shade
// joined by compacted queues in HBM (one atomic per wave: ballot + popcount +
// mbcnt):
//   trace  traces the extension queue (closest hit) and the shadow queue (any
//          hit, applying visible light samples to their paths);
//   regen  finishes the paths that ended in the previous shade (their last
//          shadow ray was just traced), averages the sample into the pixel
//          (incrementalAvg, in sample order per pixel) and queues the camera
//          ray of the pixel's next sample;
//   shade  runs one PathTracer vertex per extension hit and queues the next
//          extension ray, the light sample's shadow ray, or the path's end.
// Slots start new samples as soon as their previous sample ends, so every
// iteration traces a full queue until the last samples drain, instead of
// paying each bounce's slowest rays once per sample pass.
//
// Per-slot state (48 bytes):
//   sA = (T.xyz, flags)   flags: bit 0 found, bit 1 ends, bits 8..15 depth,
//                         bits 16..31 sample index
//   sB = (acc_all.xyz, pixel id)
//   sC = (acc_prefix.xyz, 0)
// with the chain semantics of mrt_path.hpp (PathTracer.cpp:25-164): the
// shadow result adds T_d * Ld_d and records the first vertex with
// any(Ld > 0); regen picks acc_prefix when the chain ended on an emitter.
// Valid for the PathTracer with one light sample per vertex and no
// reflection/transmission materials (the host checks; other cases use the
// megakernel).
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "mrt_device_util.hpp"
#include "mrt_wavefront.hpp"

namespace mrt {

constexpr uint32_t kFound = 1u, kEnds = 2u;
constexpr uint32_t kDeadEntry = 0xFFFFFFFFu;

__device__ __forceinline__ uint32_t flags_depth(uint32_t f) { return (f >> 8) & 0xFFu; }
__device__ __forceinline__ uint32_t flags_sample(uint32_t f) { return f >> 16; }
__device__ __forceinline__ uint32_t with_depth(uint32_t f, uint32_t d) { return (f & ~0xFF00u) | (d << 8); }
__device__ __forceinline__ uint32_t new_flags(uint32_t sample) { return (sample << 16) | (1u << 8); }

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

// Reserves `want[k] ? 1 : 0` entries of queue k per thread with one atomic
// per queue and workgroup: the waves' counts meet in LDS, K threads reserve
// the workgroup's entries, and each wave takes its range after the waves
// before it.  Same-address atomics serialise device-wide (about 10 ns each),
// so one per wave made the shading kernels wait on their queue counters most
// of their time.  Every thread of the workgroup calls it, the same number of
// times.  Returns each thread's entries in out (valid where want is true).
template <int K>
__device__ __forceinline__ void block_append(uint32_t* const (&counter)[K], const bool (&want)[K],
                                             uint32_t (&out)[K]) {
    constexpr uint32_t kWaves = kBlock / 64u;
    __shared__ uint32_t tot[K][kWaves];
    __shared__ uint32_t base[K];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long mask[K];
#pragma unroll
    for (int k = 0; k < K; ++k) mask[k] = __ballot(want[k]);
    __syncthreads();  // the previous call's readers are done with tot / base
    if (lane == 0u) {
#pragma unroll
        for (int k = 0; k < K; ++k) tot[k][wave] = static_cast<uint32_t>(__popcll(mask[k]));
    }
    __syncthreads();
    if (threadIdx.x < static_cast<uint32_t>(K)) {
        const uint32_t k = threadIdx.x;
        uint32_t sum = 0;
        for (uint32_t w = 0; w < kWaves; ++w) sum += tot[k][w];
        base[k] = sum ? atomicAdd(counter[k], sum) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        uint32_t off = base[k];
        for (uint32_t w = 0; w < wave; ++w) off += tot[k][w];
        out[k] = off + static_cast<uint32_t>(__popcll(mask[k] & below));
    }
}

// The ray counters of a workgroup, summed in LDS and added with one atomic
// each (every thread of the workgroup calls it).
__device__ __forceinline__ void block_flush_rays(const RayCounts& rc, unsigned long long* counters) {
    __shared__ uint32_t part[4][kBlock / 64u];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t v[4] = {wave_sum(rc.camera), wave_sum(rc.extension), wave_sum(rc.shadow), wave_sum(rc.specular)};
    __syncthreads();
    if (lane == 0u)
        for (int k = 0; k < 4; ++k) part[k][wave] = v[k];
    __syncthreads();
    if (threadIdx.x < 4u) {
        unsigned long long sum = 0;
        for (uint32_t w = 0; w < kBlock / 64u; ++w) sum += part[threadIdx.x][w];
        if (sum) atomicAdd(counters + threadIdx.x, sum);
    }
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

// Writes extension-queue entry `qe`, and with `keyed` its sort key when
// reordering (entries that skip the reorder, the first camera rays, have
// none: their histogram counts would shift every scattered position).
__device__ __forceinline__ void put_ray(const WfBuffers& b, float4* queue, uint32_t qe, f3 o, f3 d, uint32_t excl,
                                        uint32_t slot, bool keyed = true) {
    queue[2u * qe] = make_float4(o.x, o.y, o.z, d.x);
    queue[2u * qe + 1u] = make_float4(d.y, d.z, __uint_as_float(excl), __uint_as_float(slot));
    if (keyed && b.sort) {
        const uint32_t key = ray_key(b, o, d);
        b.keys[qe] = key;
        atomicAdd(b.hist + key, 1u);
    }
}

// Path state of a new sample of slot `slot` (pixel `pid`).
__device__ __forceinline__ void reset_path(const WfBuffers& b, uint32_t slot, uint32_t pid, uint32_t sample) {
    b.sA[slot] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(new_flags(sample)));
    b.sB[slot] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(pid));
    b.sC[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ------------------------------------------------------------- generate --
// Camera rays of the first sample of every covered slot of the group.
__global__ __launch_bounds__(kBlock) void k_wf_generate(RenderParams p, uint32_t sample, WfBuffers b, float4* queue,
                                                        uint32_t* count) {
    const uint32_t logical = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t local = logical * kBlock + threadIdx.x;
    const uint32_t slot = b.slot_base + local;
    bool live = false;
    uint32_t x = 0, y = 0;
    const uint32_t lane = slot / b.pixel_slots;  // sample lane of a sample-parallel frame
    const uint32_t s0 = sample + lane;
    if (local < b.slot_count) {
        slot_pixel(p, slot - lane * b.pixel_slots, &x, &y);
        live = pixel_covered(p, x, y) && s0 < p.spp;
    }
    f3 o{}, d{};
    if (live) {
        const uint32_t pid = y * p.width + x;
        primary_ray(p, x, y, pid, s0, &o, &d);
        reset_path(b, slot, pid, s0);
    }
    const uint32_t qe = wave_append(count, live);
    if (live) put_ray(b, queue, qe, o, d, MRT_ID_NONE, slot, false);
}

// Camera rays of the first sample of job `job` on pool waves [pool_wave,
// pool_wave + n_waves), which become the job's waves [job_wave, ...); records
// each wave's owner and the job's table entry (identical from every pipeline
// the job runs on) for the kernels that follow on this pipeline.
__global__ __launch_bounds__(kBlock) void k_wf_admit(BatchDev bd, JobDev J, uint32_t job, uint32_t pool_wave,
                                                     uint32_t job_wave, uint32_t n_waves, uint32_t first,
                                                     WfBuffers b, float4* queue, uint32_t* count) {
    const uint32_t wl = blockIdx.x * (kBlock / 64u) + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (blockIdx.x == 0u) {
        constexpr uint32_t kWords = sizeof(JobDev) / 4u;
        static_assert(sizeof(JobDev) % 4u == 0u, "job entry is copied as words");
        const uint32_t* src = reinterpret_cast<const uint32_t*>(&J);
        uint32_t* dst = reinterpret_cast<uint32_t*>(const_cast<JobDev*>(bd.jobs) + job);
        for (uint32_t k = threadIdx.x; k < kWords; k += kBlock) dst[k] = src[k];
        if (first && threadIdx.x == 0u) bd.finished[job] = 0u;
    }
    bool live = false;
    uint32_t slot = 0;
    f3 o{}, d{};
    if (wl < n_waves) {
        const uint32_t wv = pool_wave + wl;
        if (lane == 0u) bd.waves[wv] = make_uint2(job, job_wave + wl);
        slot = (wv << 6) | lane;
        // Job slot -> (lane k, pixel slot ps): sample first_sample + k.
        const uint32_t local = ((job_wave + wl) << 6) | lane;
        const uint32_t k = J.rad ? local / J.pixel_slots : 0u;
        const uint32_t ps = local - k * J.pixel_slots;
        uint32_t x, y;
        slot_pixel(J.p, J.rad ? ps : local, &x, &y);
        live = x < 16u * J.p.bsx && y < 16u * J.p.bsy;
        if (live) {
            const uint32_t pid = y * J.p.width + x;
            primary_ray(J.p, x, y, pid, J.first_sample + k, &o, &d);
            reset_path(b, slot, pid, J.first_sample + k);
        }
    }
    const uint32_t qe = wave_append(count, live);
    if (live) put_ray(b, queue, qe, o, d, MRT_ID_NONE, slot, false);
}

// Adds, per job, the lanes that finished their slot's last sample to
// bd.finished (one atomic per job present in the wave, after a fence that
// makes the lanes' bitmap and accumulation writes visible first).
__device__ __forceinline__ void count_finished(const BatchDev& bd, uint32_t job, bool last) {
    unsigned long long m = __ballot(last);
    if (m == 0ull) return;
    __threadfence();
    const uint32_t lane = threadIdx.x & 63u;
    while (m != 0ull) {
        const uint32_t leader = static_cast<uint32_t>(__ffsll(static_cast<long long>(m))) - 1u;
        const uint32_t lj = __shfl(job, static_cast<int>(leader), 64);
        const unsigned long long same = __ballot(last && job == lj);
        if (lane == leader) atomicAdd(bd.finished + lj, static_cast<uint32_t>(__popcll(same)));
        m &= ~same;
    }
}

// ------------------------------------------------------------- shadow --
// Applies a visible light sample to its path: acc_all += T_d * Ld_d, and the
// first vertex with any(Ld > 0) snapshots acc_prefix.
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

// --------------------------------------------------- trace (chunked) --
// Closest hit of every extension-queue entry, a static range per workgroup
// (scenes with planes or spheres, or a grid / naive accelerator).  Clears the
// counters that regen and shade fill next.
template <bool GENERIC, bool COUNT>
__global__ __launch_bounds__(kBlock) void k_wf_extend(SceneDev s, const float4* __restrict__ queue,
                                                      const uint32_t* __restrict__ count, float4* __restrict__ hits,
                                                      uint32_t* clear_q, uint32_t* clear_e, StackCfg sc,
                                                      unsigned long long* counters, uint32_t* err) {
    __shared__ uint32_t lds[kStackLds * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    if (gtid == 0) {
        *clear_q = 0u;
        *clear_e = 0u;
    }
    Stack st = make_stack(lds, sc, gtid);
    RayCounts rc{};
    uint32_t begin, end;
    block_range(*count, &begin, &end);
    for (uint32_t i = begin + threadIdx.x; i < end; i += kBlock) {
        const float4 a = queue[2u * i], c = queue[2u * i + 1u];
        if (__float_as_uint(c.w) == kDeadEntry) continue;
        const RayT r = make_ray(f3{a.x, a.y, a.z}, f3{a.w, c.x, c.y}, __float_as_uint(c.z));
        const HitT h = closest_hit<GENERIC, COUNT>(s, r, kRayLengthMax, st, rc.ts);
        hits[i] = make_float4(h.t, __uint_as_float(h.id), h.u, h.v);
    }
    if (COUNT) flush_counts<true>(rc, counters, st.overflow, err);
    else if (st.overflow) atomicOr(err, 1u);
}

// Shader::shadowTrace of every queued light sample, a static range per
// workgroup.
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
        apply_light(b, e);
    }
    if (COUNT) flush_counts<true>(rc, counters, st.overflow, err);
    else flush_counts<false>(rc, counters, st.overflow, err);
}

// ------------------------------------------------- trace (persistent) --
// Refill rule: a wave fetches new queue entries when at least 32 lanes are
// idle, or when no lane has a ray left.  1M-triangle 1080p frame at 16 spp
// (8 slots per pixel): refilling at 4 / 8 / 16 / 32 / 48 idle lanes gave
// 783 / 800 / 811 / 815 / 791 Mrays/s (a refill is a ballot, a fetch and a
// ray setup for the whole wave, so waiting for more idle lanes amortises it
// until the idle lanes cost more than it).
__device__ __forceinline__ bool wave_wants_refill(bool need, bool has, uint32_t refill_lanes) {
    const unsigned long long needMask = __ballot(need);
    return needMask != 0ull && (static_cast<uint32_t>(__popcll(needMask)) >= refill_lanes || __ballot(has) == 0ull);
}

// The leaf test over the compressed tree's leaf records (mrt_cwt.hpp): one
// record per step, its exact leaf box tested first (lanes with a finite
// inverse direction reached the leaf through quantized boxes; the others
// through the pair tree's exact boxes, so they skip it), then its one or two
// triangles as walk_test tests them.  A box the ray misses drops the leaf.
template <bool COUNT>
__device__ __forceinline__ bool cw_test(const SceneDev& s, Walk& w, bool shadow, bool finite, TraceStats& ts) {
    const uint32_t rec = w.cur & kSlotMask;
    const uint32_t cnt = w.cur >> kSlotBits;
    const float4* p = s.cw_recs + 8u * rec;
    const float4 b0 = p[0], b1 = p[1];
    const float4 a0 = p[2], a1 = p[3], a2 = p[4];
    const float4 c0 = p[5], c1 = p[6], c2 = p[7];
    float eb;
    const bool boxOk = !finite || (slab_finite(f3{b0.x, b0.y, b0.z}, f3{b0.w, b1.x, b1.y}, w.r.o, w.r.inv, &eb) &&
                                   eb <= cull_limit(w.h.t));
    if (!boxOk) {
        w.cur = 0u;
        walk_pull(w);
        return false;
    }
    const TriG ga{f3{a0.x, a0.y, a0.z}, f3{a0.w, a1.x, a1.y}, f3{a1.z, a1.w, a2.x}};
    const TriG gb{f3{c0.x, c0.y, c0.z}, f3{c0.w, c1.x, c1.y}, f3{c1.z, c1.w, c2.x}};
    const uint32_t ida = __float_as_uint(a2.y), idb = __float_as_uint(c2.y);
    const bool two = cnt > 1u;
    if (COUNT) ts.prims[shadow ? 2 : (rec >= s.walk_light_slot ? 1 : 0)] += two ? 2u : 1u;
    float t, u, v;
    bool hit = tri_test_flat(ga, w.r, &t, &u, &v) && ida != w.r.excl;
    bool better = !shadow && hit && key_better(t, ida, w.h);
    bool blocked = shadow && hit && t > kEpsilon && t < w.h.t;
    w.h.t = better ? t : w.h.t;
    w.h.id = better ? ida : w.h.id;
    w.h.u = better ? u : w.h.u;
    w.h.v = better ? v : w.h.v;
    hit = two && tri_test_flat(gb, w.r, &t, &u, &v) && idb != w.r.excl;
    better = !shadow && hit && key_better(t, idb, w.h);
    blocked = blocked || (shadow && hit && t > kEpsilon && t < w.h.t);
    w.h.t = better ? t : w.h.t;
    w.h.id = better ? idb : w.h.id;
    w.h.u = better ? u : w.h.u;
    w.h.v = better ? v : w.h.v;
    // Next record of a longer leaf (count - 2, record + 1), else the next leaf.
    w.cur = cnt > 2u ? ((cnt - 2u) << kSlotBits) | (rec + 1u) : 0u;
    walk_pull(w);
    if (blocked) walk_clear(w);
    return blocked;
}

// One unit of walk work per lane: the test of one leaf record (when a leaf
// is pending) and a node visit (when none is queued).  Lanes with a finite
// inverse direction walk the compressed wide tree, the others the pair tree
// with the reference slab test; both test the leaf records.  Returns true
// when a shadow lane found a blocker.
// OVERLAP: both loads of the step go out before either result is used (the
// node a lane visits does not depend on its test), so a ray's two memory
// round trips per step overlap; the node's registers are then held across
// the test (80 instead of 67 registers, 6 instead of 7 waves per SIMD).  That
// wins when a launch holds too few rays to fill the chip (a ray's walk
// latency sets the launch time) and loses on full launches, so both are built
// and the host picks by queue size.
template <bool COUNT, bool OVERLAP>
__device__ __forceinline__ bool walk_work(const SceneDev& s, Walk& w, bool has, bool shadow, bool finite, Stack& st,
                                          TraceStats& ts) {
    const bool canVisit = has && walk_can_visit(w);
    const bool canTest = has && w.cur != 0u;
    bool occluded = false;
    CwNode cn{};
    if (OVERLAP && canVisit && finite) cn = load_cw(s, w.node);
    if (canTest) occluded = cw_test<COUNT>(s, w, shadow, finite, ts);
    if (canVisit && !occluded) {
        const float lim = cull_limit(w.h.t);
        const int kind = shadow ? 2 : 0;
        if (finite) cw_visit<COUNT>(s, w, OVERLAP ? cn : load_cw(s, w.node), lim, st, ts, kind);
        else walk_visit<false, COUNT>(s, w, lim, st, ts, kind);
    }
    return occluded;
}

// Wave-private run of trace-queue entries [next, end): a wave reserves
// `chunk` (>= 64) entries with one atomic and hands them to its lanes as they
// go idle.  Same-address atomics serialise device-wide (about 10 ns each), so
// one atomic per refill of a few lanes would cap the whole trace at ~100 M
// refills per second.
struct FetchRun {
    uint32_t next, end;  // equal in every lane of the wave
    bool drained;        // a run reached the queue's end: reserve no more
};

// Next queue entry for each lane that wants one (n entries in all);
// kDeadEntry once the queue is used up.
__device__ __forceinline__ uint32_t fetch_entry(uint32_t* cursor, uint32_t n, bool want, FetchRun& run,
                                                uint32_t chunk) {
    const unsigned long long mask = __ballot(want);
    const uint32_t total = static_cast<uint32_t>(__popcll(mask));
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t rank = static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
    const uint32_t avail = run.end - run.next;
    uint32_t e = run.next + rank;
    if (total <= avail) {
        run.next += total;
    } else if (!run.drained) {
        // The run's rest goes to the first `avail` lanes, a new run to the
        // others (total <= 64 <= chunk).
        const uint32_t leader = static_cast<uint32_t>(__ffsll(static_cast<long long>(mask))) - 1u;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(cursor, chunk);
        base = __shfl(base, static_cast<int>(leader), 64);
        if (rank >= avail) e = base + (rank - avail);
        run.next = base + (total - avail);
        run.end = base + chunk;
        run.drained = base + chunk >= n;
    } else {
        if (rank >= avail) e = kDeadEntry;
        run.next = run.end;
    }
    return e < n ? e : kDeadEntry;
}

// Persistent trace of the extension queue followed by the shadow queue, for
// BVH scenes of triangles and lights: a lane takes the next entry the moment
// its ray is done, so a wave never waits for its slowest ray.  Clears the
// counters that regen and shade fill next.
// MULTI: the rays belong to the jobs of a batch and start at their scenes'
// roots in the walk arena `s`.
template <bool COUNT, bool MULTI, bool OVERLAP>
__global__ __launch_bounds__(kBlock) void k_wf_trace_p(SceneDev s, const float4* __restrict__ queue,
                                                       const uint32_t* __restrict__ count_q,
                                                       const uint32_t* __restrict__ count_s, float4* __restrict__ hits,
                                                       WfBuffers b, uint32_t* clear_q, uint32_t* clear_e,
                                                       uint32_t* fetch, StackCfg sc, unsigned long long* counters,
                                                       uint32_t* err, BatchDev bd) {
    __shared__ uint32_t lds[kStackLds * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    if (gtid == 0) {
        *clear_q = 0u;
        *clear_e = 0u;
    }
    Stack st = make_stack(lds, sc, gtid);
    RayCounts rc{};
    const uint32_t nq = *count_q;
    const uint32_t n = nq + *count_s;
    Walk w;
    walk_clear(w);
    w.h.t = 0.0f;
    uint32_t idx = 0;
    bool has = false, exhausted = false, finite = true, shadow = false;
    FetchRun run{0u, 0u, false};
    while (true) {
        const bool need = !has && !exhausted;
        if (wave_wants_refill(need, has, b.refill_lanes)) {
            const uint32_t e = fetch_entry(fetch, n, need, run, b.fetch_chunk);
            if (need) {
                if (e < nq) {
                    const float4 a = queue[2u * e], c = queue[2u * e + 1u];
                    if (__float_as_uint(c.w) != kDeadEntry) {
                        idx = e;
                        shadow = false;
                        w.r = make_ray(f3{a.x, a.y, a.z}, f3{a.w, c.x, c.y}, __float_as_uint(c.z));
                        finite = inv_finite(w.r);
                        uint32_t root = kWalkClosestRoot;
                        if (MULTI) {
                            const JobDev& J = bd.jobs[bd.waves[__float_as_uint(c.w) >> 6].x];
                            root = finite ? J.wide_root : J.walk_root;
                        }
                        walk_begin(w, root, kRayLengthMax, st);
                        has = true;
                    }
                } else if (e < n) {
                    idx = e - nq;
                    const float4* sh = b.shadow + 4u * idx;
                    const float4 a = sh[0], c = sh[1];
                    shadow = true;
                    w.r = make_ray(f3{a.x, a.y, a.z}, f3{a.w, c.x, c.y}, __float_as_uint(c.w));
                    finite = inv_finite(w.r);
                    uint32_t root = kWalkAnyRoot;
                    if (MULTI) {
                        const JobDev& J = bd.jobs[__float_as_uint(sh[3].w)];
                        root = (finite ? J.wide_root : J.walk_root) + 1u;
                    }
                    walk_begin(w, root, c.z, st);
                    ++rc.shadow;
                    has = true;
                } else {
                    exhausted = true;
                }
            }
        }
        if (__ballot(has || !exhausted) == 0ull) break;
        if (COUNT) {
            ++rc.ts.lane_slots;
            if (has) ++rc.ts.lane_iters;
        }
        const bool occluded = walk_work<COUNT, OVERLAP>(s, w, has, shadow, finite, st, rc.ts);
        if (has && walk_done(w)) {
            if (!shadow) hits[idx] = make_float4(w.h.t, __uint_as_float(w.h.id), w.h.u, w.h.v);
            else if (!occluded) apply_light(b, b.shadow + 4u * idx);
            has = false;
        }
    }
    if (COUNT) flush_counts<true>(rc, counters, st.overflow, err);
    else flush_counts<false>(rc, counters, st.overflow, err);
}

// ------------------------------------------- trace (persistent, grid) --
// Persistent trace of both queues for RegularGrid scenes of triangles and
// lights: the refilling lanes of k_wf_trace_p over a resumable 3D-DDA.  A
// lane's unit of work is up to kGridBatch primitive tests of its current
// cell's list or one DDA step to the next cell (RegularGrid<T>::trace / shadowTrace,
// RegularGrid.hpp:157-465, in their order).  Closest hits: phase 1 tests the
// cell's list until a primitive hits; the rest of that cell's list is then
// tested in phase 2 (the reference's phase 2 re-tests the whole list, but a
// primitive already tested misses again or returns t equal to the kept one,
// which the strict t < length_ rejects); phase 2 steps on only while the
// nearest t is not below the next wall.  The area lights are tested after the
// walk (Scene::traceLights after the accelerators, Shader.cpp:144-149).  Any
// hits return at the first blocker.
// Primitives tested per step: their loads go out together (1M-soup 1080p
// 16-sample grid frame with 1 / 2 / 4: 141.7 / 181.4 / 190.6 Mrays/s).
constexpr uint32_t kGridBatch = 4;

struct GridWalk {
    Dda g;
    uint32_t pos, end;  // untested part of the current cell's list
    bool phase2, done;
};

__device__ __forceinline__ void grid_walk_begin(const Accel& a, const RayT& r, GridWalk& w) {
    dda_setup(a, r, w.g);
    const uint32_t c = dda_cell(a, w.g);
    w.pos = a.cell_off[c];
    w.end = a.cell_off[c + 1u];
    w.phase2 = false;
    w.done = false;
}

template <bool COUNT>
__global__ __launch_bounds__(kBlock) void k_wf_trace_grid_p(SceneDev s, const float4* __restrict__ queue,
                                                            const uint32_t* __restrict__ count_q,
                                                            const uint32_t* __restrict__ count_s,
                                                            float4* __restrict__ hits, WfBuffers b, uint32_t* clear_q,
                                                            uint32_t* clear_e, uint32_t* fetch, StackCfg sc,
                                                            unsigned long long* counters, uint32_t* err) {
    __shared__ uint32_t lds[kStackLds * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    if (gtid == 0) {
        *clear_q = 0u;
        *clear_e = 0u;
    }
    Stack st = make_stack(lds, sc, gtid);
    RayCounts rc{};
    const Accel& a = s.acc[MRT_CAT_TRIANGLE];
    const uint32_t nq = *count_q;
    const uint32_t n = nq + *count_s;
    RayT r{};
    HitT h{};
    GridWalk w{};
    uint32_t idx = 0;
    bool has = false, exhausted = false, shadow = false;
    FetchRun run{0u, 0u, false};
    while (true) {
        const bool need = !has && !exhausted;
        if (wave_wants_refill(need, has, b.refill_lanes)) {
            const uint32_t e = fetch_entry(fetch, n, need, run, b.fetch_chunk);
            if (need) {
                if (e < nq) {
                    const float4 qa = queue[2u * e], qc = queue[2u * e + 1u];
                    if (__float_as_uint(qc.w) != kDeadEntry) {
                        idx = e;
                        shadow = false;
                        r = make_ray(f3{qa.x, qa.y, qa.z}, f3{qa.w, qc.x, qc.y}, __float_as_uint(qc.z));
                        h = HitT{kRayLengthMax, MRT_ID_NONE, 0.0f, 0.0f};
                        grid_walk_begin(a, r, w);
                        has = true;
                    }
                } else if (e < n) {
                    idx = e - nq;
                    const float4* sh = b.shadow + 4u * idx;
                    const float4 qa = sh[0], qc = sh[1];
                    shadow = true;
                    r = make_ray(f3{qa.x, qa.y, qa.z}, f3{qa.w, qc.x, qc.y}, __float_as_uint(qc.w));
                    h = HitT{qc.z, MRT_ID_NONE, 0.0f, 0.0f};  // h.t holds the shadow distance
                    grid_walk_begin(a, r, w);
                    ++rc.shadow;
                    has = true;
                } else {
                    exhausted = true;
                }
            }
        }
        if (__ballot(has || !exhausted) == 0ull) break;
        if (COUNT) {
            ++rc.ts.lane_slots;
            if (has) ++rc.ts.lane_iters;
        }
        bool occluded = false;
        if (has && w.pos < w.end) {
            // Up to kGridBatch primitives of the cell's list per step, all
            // loaded before any is tested and tested in list order (the strict
            // t < best rule keeps the first of equal hits, as one at a time).
            const uint32_t m = min(w.end - w.pos, kGridBatch);
            uint32_t pi[kGridBatch];
#pragma unroll
            for (uint32_t k = 0; k < kGridBatch; ++k) {
                pi[k] = 0u;
                if (k < m) pi[k] = a.cell_ref[w.pos + k];
            }
            w.pos += m;
            TriG g[kGridBatch];
#pragma unroll
            for (uint32_t k = 0; k < kGridBatch; ++k) g[k] = load_tri(s.tri_geo, k < m ? pi[k] : pi[0]);
            if (COUNT) rc.ts.prims[shadow ? 2 : 0] += m;
            if (shadow) {
#pragma unroll
                for (uint32_t k = 0; k < kGridBatch; ++k)
                    occluded = occluded ||
                               (k < m && MRT_MAKE_ID(MRT_CAT_TRIANGLE, pi[k]) != r.excl && tri_any(g[k], r, h.t));
                w.done = occluded;
            } else {
                bool hit = false;
#pragma unroll
                for (uint32_t k = 0; k < kGridBatch; ++k)
                    if (k < m) hit = tri_closest_strict(g[k], pi[k], r, h) || hit;
                if (hit) w.phase2 = true;
            }
        } else if (has) {
            if (COUNT) ++rc.ts.pairs[shadow ? 2 : 0];
            if (!dda_step(w.g, !shadow && w.phase2, h.t)) {
                w.done = true;
            } else {
                const uint32_t c = dda_cell(a, w.g);
                w.pos = a.cell_off[c];
                w.end = a.cell_off[c + 1u];
            }
        }
        if (has && w.done) {
            if (!shadow) {
                lights_closest<COUNT>(s, r, h, st, rc.ts);
                hits[idx] = make_float4(h.t, __uint_as_float(h.id), h.u, h.v);
            } else if (!occluded) {
                apply_light(b, b.shadow + 4u * idx);
            }
            has = false;
        }
    }
    if (COUNT) flush_counts<true>(rc, counters, st.overflow, err);
    else flush_counts<false>(rc, counters, st.overflow, err);
}

// ----------------------------------------------- persistent megakernel --
// One launch renders the whole frame: every lane takes work items (a pixel
// slot and a sample lane k, whose samples first + k, first + k + lanes, ...
// it renders in order) and keeps its path in registers, walking the combined
// walk tree one unit per iteration like the persistent trace.  A lane whose
// walk ended needs service (shade the vertex and start the shadow or
// extension walk, apply a light sample, finish a sample and start the next,
// or take a new item); a wave runs the service step once `service_lanes` of
// its lanes need it, or when no lane is walking, so the shading code runs for
// many lanes at once instead of for each lane whose walk just ended.  Each
// path's operations and their order are those of PathTracer::shade as the
// wavefront kernels run it (shade, then its shadow ray, then its extension
// ray), so frames are identical to theirs.

enum : uint32_t { kMegaIdle = 0, kMegaExt = 1, kMegaShadow = 2 };

// A lane's item, path and pending rays between its walks.
struct MegaState {
    f3 T, accAll, accPrefix;  // path state (sA, sB, sC of the wavefront)
    f3 vO, shD, extD, contrib;  // the shaded vertex's shadow and extension rays
    float shDist;
    uint32_t vExcl, ps, x, y, pid, sample, depth;
    bool found, ends, ldPos, pendExt;
};

// MEM: the state lives in device memory (7 float4 per lane) between
// services instead of registers.
__device__ __forceinline__ void mega_store(float4* q, uint32_t stride, const MegaState& m) {
    const uint32_t fl = (m.found ? 1u : 0u) | (m.ends ? 2u : 0u) | (m.ldPos ? 4u : 0u) | (m.pendExt ? 8u : 0u) |
                        (m.depth << 8);
    q[0] = make_float4(m.T.x, m.T.y, m.T.z, __uint_as_float(fl));
    q[stride] = make_float4(m.accAll.x, m.accAll.y, m.accAll.z, __uint_as_float(m.pid));
    q[2u * stride] = make_float4(m.accPrefix.x, m.accPrefix.y, m.accPrefix.z, __uint_as_float(m.sample));
    q[3u * stride] = make_float4(m.contrib.x, m.contrib.y, m.contrib.z, m.shDist);
    q[4u * stride] = make_float4(m.shD.x, m.shD.y, m.shD.z, __uint_as_float(m.vExcl));
    q[5u * stride] = make_float4(m.extD.x, m.extD.y, m.extD.z, __uint_as_float(m.ps));
    q[6u * stride] = make_float4(m.vO.x, m.vO.y, m.vO.z, __uint_as_float(m.x | (m.y << 16)));
}
__device__ __forceinline__ void mega_load(const float4* q, uint32_t stride, MegaState& m) {
    const float4 a = q[0], b = q[stride], c = q[2u * stride], d = q[3u * stride];
    const float4 e = q[4u * stride], f = q[5u * stride], g = q[6u * stride];
    const uint32_t fl = __float_as_uint(a.w);
    m.T = f3{a.x, a.y, a.z};
    m.found = fl & 1u;
    m.ends = fl & 2u;
    m.ldPos = fl & 4u;
    m.pendExt = fl & 8u;
    m.depth = fl >> 8;
    m.accAll = f3{b.x, b.y, b.z};
    m.pid = __float_as_uint(b.w);
    m.accPrefix = f3{c.x, c.y, c.z};
    m.sample = __float_as_uint(c.w);
    m.contrib = f3{d.x, d.y, d.z};
    m.shDist = d.w;
    m.shD = f3{e.x, e.y, e.z};
    m.vExcl = __float_as_uint(e.w);
    m.extD = f3{f.x, f.y, f.z};
    m.ps = __float_as_uint(f.w);
    m.vO = f3{g.x, g.y, g.z};
    const uint32_t xy = __float_as_uint(g.w);
    m.x = xy & 0xFFFFu;
    m.y = xy >> 16;
}

__device__ __forceinline__ void mega_walk(Walk& w, bool& finite, uint32_t& mode, bool& occluded, f3 o, f3 d,
                                          uint32_t excl, bool shadowRay, float dist, Stack& st) {
    w.r = make_ray(o, d, excl);
    finite = inv_finite(w.r);
    walk_begin(w, shadowRay ? kWalkAnyRoot : kWalkClosestRoot, shadowRay ? dist : kRayLengthMax, st);
    mode = shadowRay ? kMegaShadow : kMegaExt;
    occluded = false;
}

__device__ __forceinline__ void mega_start_sample(const RenderParams& p, MegaState& m, Walk& w, bool& finite,
                                                  uint32_t& mode, bool& occluded, Stack& st) {
    f3 o, d;
    primary_ray(p, m.x, m.y, m.pid, m.sample, &o, &d);
    m.T = f3{1.0f, 1.0f, 1.0f};
    m.accAll = m.accPrefix = f3{0.0f, 0.0f, 0.0f};
    m.found = m.ends = false;
    m.depth = 1u;
    mega_walk(w, finite, mode, occluded, o, d, MRT_ID_NONE, false, 0.0f, st);
}

// The service of a lane whose walk ended: applies a light sample and starts
// the pending extension walk, or shades the vertex the extension (or camera)
// ray reached (PathTracer::shade, PathTracer.cpp:25-133, as k_wf_shade runs
// it) and starts its shadow or extension walk, or finishes the sample
// (k_wf_regen's result selection and running mean, Renderer.cpp:101-104)
// and starts the lane's next one.
__device__ __forceinline__ void mega_service(const SceneDev& s, const MegaArgs& A, MegaState& m, Walk& w,
                                             bool& finite, uint32_t& mode, bool& occluded, Stack& st, RayCounts& rc,
                                             uint32_t& finished) {
    const RenderParams& p = A.p;
    bool finish = false;
    if (mode == kMegaShadow) {
        if (!occluded) {
            m.accAll = add(m.accAll, m.contrib);
            if (!m.found && m.ldPos) {
                m.found = true;
                m.accPrefix = m.accAll;
            }
        }
        if (m.pendExt) mega_walk(w, finite, mode, occluded, m.vO, m.extD, m.vExcl, false, 0.0f, st);
        else finish = true;
    } else {
        if (m.depth == 1u) ++rc.camera;
        else ++rc.extension;
        bool wantShadow = false;
        m.pendExt = false;
        if (w.h.id != MRT_ID_NONE) {
            const Surf sf = surface(s, w.r, w.h);
            const Shading sh = shading_of(s, sf);
            const DrawCtx dc{m.pid, m.sample, 0u};
            const int32_t depth = static_cast<int32_t>(m.depth);
            if (sh.emitter) {
                m.accAll = add(m.accAll, mul(m.T, sh.le));
                m.ends = true;
            } else if (any_gt0(sh.kd)) {
                if (s.n_lights > 0u) {
                    f3 lc;
                    if (light_sample(s, p, sf, depth, dc, 0u, &m.shD, &m.shDist, &lc)) {
                        const f3 Ld = divs(mul(lc, sh.kd), static_cast<float>(p.samples_light));
                        m.contrib = mul(m.T, Ld);
                        m.ldPos = any_gt0(Ld);
                        wantShadow = true;
                    }
                }
                bool cont = depth <= p.depth_min;
                if (!cont) cont = draw_at(p, p.s_rr, dc, kSlotRR, 0, depth) > 0.5f;
                if (cont && depth + 1 <= p.depth_max) {
                    m.extD = cosine_hemisphere(p, sf.normal, dc, depth);
                    f3 Tn = mul(m.T, sh.kd);
                    if (depth > p.depth_min) Tn = divs(Tn, 0.5f * 0.5f);
                    m.T = Tn;
                    ++m.depth;
                    m.pendExt = true;
                }
                m.vO = sf.point;
                m.vExcl = sf.excl;
            }
        }
        if (wantShadow) {
            ++rc.shadow;
            mega_walk(w, finite, mode, occluded, m.vO, m.shD, m.vExcl, true, m.shDist, st);
        } else if (m.pendExt) {
            mega_walk(w, finite, mode, occluded, m.vO, m.extD, m.vExcl, false, 0.0f, st);
        } else {
            finish = true;
        }
    }
    if (finish) {
        const f3 c = (m.ends && m.found) ? m.accPrefix : m.accAll;
        if (A.rad) {
            float* r = A.rad + 3ull * ((m.sample - A.first) * static_cast<uint64_t>(A.pixel_slots) + m.ps);
            r[0] = c.x;
            r[1] = c.y;
            r[2] = c.z;
        } else {
            const uint32_t bi = m.y * p.real_width + m.x;
            A.bitmap[bi] = incremental_avg(c, A.bitmap[bi], m.sample + 1u);
            if (A.accum) {
                A.accum[3u * m.pid + 0u] += c.x;
                A.accum[3u * m.pid + 1u] += c.y;
                A.accum[3u * m.pid + 2u] += c.z;
            }
        }
        ++finished;
        m.sample += A.lanes;
        if (m.sample < A.spp) {
            mega_start_sample(p, m, w, finite, mode, occluded, st);
        } else {
            mode = kMegaIdle;
            walk_clear(w);
        }
    }
}

template <bool COUNT, bool MEM>
// Four waves per SIMD (at most 128 VGPRs): the walk waits on dependent node
// loads, and a fourth wave hides more of that wait than the spills of the
// shading step cost (1M-triangle 1080p 16-spp frame: 131 -> 122 ms).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_mega_p(SceneDev s, MegaArgs A, StackCfg sc, unsigned long long* counters,
                                                   uint32_t* err) {
    __shared__ uint32_t lds[kStackLds * kBlock];
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t lanesAll = gridDim.x * kBlock;
    Stack st = make_stack(lds, sc, gtid);
    RayCounts rc{};
    const uint32_t nItems = A.lanes * A.pixel_slots;
    MegaState m{};
    float4* mq = MEM ? A.state + gtid : nullptr;
    Walk w;
    walk_clear(w);
    w.h.t = 0.0f;
    uint32_t mode = kMegaIdle;
    bool finite = true, occluded = false, exhausted = false;
    uint32_t finished = 0;
    FetchRun run{0u, 0u, false};
    while (true) {
        const bool walking = mode != kMegaIdle && !walk_done(w);
        const bool needs = (mode != kMegaIdle && !walking) || (mode == kMegaIdle && !exhausted);
        if (__ballot(walking || needs) == 0ull) break;
        const unsigned long long svc = __ballot(needs);
        if (svc != 0ull && (static_cast<uint32_t>(__popcll(svc)) >= A.service_lanes || __ballot(walking) == 0ull)) {
            if (needs && mode != kMegaIdle) {
                if (MEM) {
                    MegaState ml;
                    mega_load(mq, lanesAll, ml);
                    mega_service(s, A, ml, w, finite, mode, occluded, st, rc, finished);
                    if (mode != kMegaIdle) mega_store(mq, lanesAll, ml);
                } else {
                    mega_service(s, A, m, w, finite, mode, occluded, st, rc, finished);
                }
            }
            // New items for idle lanes (every lane of the wave takes part in
            // the fetch's ballot).
            const bool want = mode == kMegaIdle && !exhausted;
            const uint32_t e = fetch_entry(A.fetch, nItems, want, run, A.fetch_chunk);
            if (want) {
                if (e == kDeadEntry) {
                    exhausted = true;
                } else {
                    MegaState mf{};
                    MegaState& mm = MEM ? mf : m;
                    const uint32_t k = e / A.pixel_slots;
                    mm.ps = e - k * A.pixel_slots;
                    slot_pixel(A.p, mm.ps, &mm.x, &mm.y);
                    mm.sample = A.first + k;
                    if (pixel_covered(A.p, mm.x, mm.y) && mm.sample < A.spp) {
                        mm.pid = mm.y * A.p.width + mm.x;
                        mega_start_sample(A.p, mm, w, finite, mode, occluded, st);
                        if (MEM) mega_store(mq, lanesAll, mm);
                    }
                }
            }
            continue;
        }
        if (COUNT) {
            ++rc.ts.lane_slots;
            if (walking) ++rc.ts.lane_iters;
        }
        if (walk_work<COUNT, false>(s, w, walking, mode == kMegaShadow, finite, st, rc.ts)) occluded = true;
    }
    const uint32_t total = wave_sum(finished);
    if ((threadIdx.x & 63u) == 0u && total) atomicAdd(A.done, total);
    if (COUNT) flush_counts<true>(rc, counters, st.overflow, err);
    else flush_counts<false>(rc, counters, st.overflow, err);
}

uint32_t mega_blocks(uint32_t num_cu) {
    static int perCu = 0;
    if (perCu == 0) {
        int a = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_mega_p<false, true>, kBlock, 0) != hipSuccess) {
            (void)hipGetLastError();
            a = 2;
        }
        perCu = std::max(1, a);
    }
    return static_cast<uint32_t>(perCu) * num_cu;
}

hipError_t launch_mega_p(const SceneDev& s, bool count, const MegaArgs& a, uint32_t blocks, const StackCfg& sc,
                         unsigned long long* counters, uint32_t* err, hipStream_t stream) {
    if (a.lanes == 0u || a.pixel_slots == 0u || (a.lanes > 1u) != (a.rad != nullptr) || a.fetch_chunk < 64u)
        return hipErrorInvalidValue;
    if (a.state) {
        if (count) k_mega_p<true, true><<<blocks, kBlock, 0, stream>>>(s, a, sc, counters, err);
        else k_mega_p<false, true><<<blocks, kBlock, 0, stream>>>(s, a, sc, counters, err);
    } else {
        if (count) k_mega_p<true, false><<<blocks, kBlock, 0, stream>>>(s, a, sc, counters, err);
        else k_mega_p<false, false><<<blocks, kBlock, 0, stream>>>(s, a, sc, counters, err);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------- regen --
// Finishes the paths that ended in the previous shade (PathTracer.cpp:144-164
// result selection, then incrementalAvg of Renderer.cpp:101-104 with the
// slot's own sample count) and starts each slot's next sample.  Clears the
// shadow counter and the trace fetch cursors for the next iteration.
// MULTI: each slot's parameters, sample count, stop flag and outputs come
// from its job, and finished jobs are counted in bd.finished.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_wf_regen(RenderParams p0, uint32_t spp0, uint32_t stop0,
                                                     const uint32_t* __restrict__ ended,
                                                     const uint32_t* __restrict__ count_e, WfBuffers b,
                                                     float4* __restrict__ queue_out, uint32_t* count_out,
                                                     uint32_t* clear_s, uint32_t* clear_fetch, uint32_t* done,
                                                     uint32_t* bitmap0, float* accum0, BatchDev bd) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *clear_s = 0u;
        *clear_fetch = 0u;
    }
    const uint32_t n = *count_e;
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t rounds = (n + stride - 1u) / stride;
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    uint32_t finished = 0;
    for (uint32_t k = 0; k < rounds; ++k) {
        const uint32_t i = gtid + k * stride;
        bool again = false, last = false;
        uint32_t slot = 0, job = 0;
        f3 o{}, d{};
        if (i < n) {
            slot = ended[i];
            uint32_t local = slot;
            if (MULTI) {
                const uint2 wv = bd.waves[slot >> 6];
                job = wv.x;
                local = (wv.y << 6) | (slot & 63u);
            }
            const JobDev* J = bd.jobs + job;  // read only when MULTI
            const RenderParams& p = MULTI ? J->p : p0;
            const uint32_t spp = MULTI ? J->spp : spp0;
            const bool stop = MULTI ? J->stop != 0u : stop0 != 0u;
            uint32_t* bitmap = MULTI ? J->bitmap : bitmap0;
            float* accum = MULTI ? J->accum : accum0;
            const float4 sa = b.sA[slot], sbv = b.sB[slot];
            const uint32_t flags = __float_as_uint(sa.w);
            f3 c = f3{sbv.x, sbv.y, sbv.z};
            if ((flags & kEnds) && (flags & kFound)) {
                const float4 scv = b.sC[slot];
                c = f3{scv.x, scv.y, scv.z};
            }
            const uint32_t pid = __float_as_uint(sbv.w);
            const uint32_t sample = flags_sample(flags);
            uint32_t step = 1u;
            uint32_t x, y;
            if (!MULTI && b.rad) {
                // Sample-parallel frame: keep the radiance for k_wf_resolve.
                const uint32_t lane = slot / b.pixel_slots;
                const uint32_t ps = slot - lane * b.pixel_slots;
                slot_pixel(p, ps, &x, &y);
                float* r = b.rad + 3ull * ((sample - b.first) * static_cast<uint64_t>(b.pixel_slots) + ps);
                r[0] = c.x;
                r[1] = c.y;
                r[2] = c.z;
                step = b.lanes;
            } else if (MULTI && J->rad) {
                // Sample-parallel job: its resolve folds the samples.
                const uint32_t lane = local / J->pixel_slots;
                const uint32_t ps = local - lane * J->pixel_slots;
                slot_pixel(p, ps, &x, &y);
                float* r = J->rad + 3ull * ((sample - J->first_sample) * static_cast<uint64_t>(J->pixel_slots) + ps);
                r[0] = c.x;
                r[1] = c.y;
                r[2] = c.z;
                step = J->lanes;
            } else {
                slot_pixel(p, local, &x, &y);
                const uint32_t bi = y * p.real_width + x;
                bitmap[bi] = incremental_avg(c, bitmap[bi], sample + 1u);
                if (accum) {
                    accum[3u * pid + 0u] += c.x;
                    accum[3u * pid + 1u] += c.y;
                    accum[3u * pid + 2u] += c.z;
                }
            }
            ++finished;
            if (!stop && sample + step < spp) {
                primary_ray(p, x, y, pid, sample + step, &o, &d);
                reset_path(b, slot, pid, sample + step);
                again = true;
            } else {
                last = true;
            }
        }
        uint32_t* const cq[1] = {count_out};
        const bool wq[1] = {again};
        uint32_t qe[1];
        block_append<1>(cq, wq, qe);
        if (again) put_ray(b, queue_out, qe[0], o, d, MRT_ID_NONE, slot);
        if (MULTI) count_finished(bd, job, last);
    }
    const uint32_t total = wave_sum(finished);
    if ((threadIdx.x & 63u) == 0u && total) atomicAdd(done, total);
}

// ------------------------------------------------------------- shade --
// One PathTracer vertex per extension-queue entry (PathTracer.cpp:25-133):
// queues the next extension ray, the light sample's shadow ray, and the
// slot when its path ends here.
// MULTI: each entry's scene and parameters come from its slot's job; a
// shadow entry records the job for the trace.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_wf_shade(SceneDev s0, RenderParams p0,
                                                     const float4* __restrict__ queue_in,
                                                     const uint32_t* __restrict__ count_in,
                                                     const float4* __restrict__ hits, WfBuffers b,
                                                     float4* __restrict__ queue_out, uint32_t* count_out,
                                                     uint32_t* count_shadow, uint32_t* __restrict__ ended,
                                                     uint32_t* count_ended, unsigned long long* counters,
                                                     BatchDev bd) {
    RayCounts rc{};
    const uint32_t n = *count_in;
    const uint32_t stride = gridDim.x * kBlock;
    // Every lane of a wave runs the same number of iterations so that the
    // wave-wide appends below see all lanes.
    const uint32_t rounds = (n + stride - 1u) / stride;
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    for (uint32_t k = 0; k < rounds; ++k) {
        const uint32_t i = gtid + k * stride;
        bool wantExt = false, wantShadow = false, wantEnd = false;
        f3 nextO{}, nextD{}, shD{}, Ld{}, Tv{};
        uint32_t excl = MRT_ID_NONE, slot = kDeadEntry, job = 0;
        float dist = 0.0f;
        if (i < n) {
            const float4 a = queue_in[2u * i], c = queue_in[2u * i + 1u];
            slot = __float_as_uint(c.w);
            if (slot != kDeadEntry) {
                if (MULTI) job = bd.waves[slot >> 6].x;
                const JobDev* J = bd.jobs + job;  // read only when MULTI
                const RenderParams& p = MULTI ? J->p : p0;
                const SceneDev& s = MULTI ? bd.scenes[J->scene] : s0;
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
                const uint32_t sample = flags_sample(flags);
                if (depth == 1u) ++rc.camera;
                else ++rc.extension;
                float4* aovEntry = nullptr;
                if (!MULTI && b.aov_s && depth == 1u) {
                    const uint32_t ps = slot % b.pixel_slots;
                    aovEntry = b.aov_s + 2ull * ((sample - b.first) * static_cast<uint64_t>(b.pixel_slots) + ps);
                    if (h.id == MRT_ID_NONE) {
                        aovEntry[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        aovEntry[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    }
                }
                if (h.id != MRT_ID_NONE) {
                    const f3 T = f3{sa.x, sa.y, sa.z};
                    const Surf sf = surface(s, r, h);
                    const Shading m = shading_of(s, sf);
                    if (aovEntry) {
                        // The first hit of the sample (the quantities of k_aov).
                        const f3 alb = m.emitter ? m.le : m.kd;
                        aovEntry[0] = make_float4(alb.x, alb.y, alb.z, h.t);
                        aovEntry[1] = make_float4(sf.normal.x, sf.normal.y, sf.normal.z, 1.0f);
                    }
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
                wantEnd = !wantExt;
            }
        }
        uint32_t* const cq[3] = {count_out, count_shadow, count_ended};
        const bool wq[3] = {wantExt, wantShadow, wantEnd};
        uint32_t q[3];
        block_append<3>(cq, wq, q);
        const uint32_t qe = q[0], qs = q[1], qd = q[2];
        if (wantExt) put_ray(b, queue_out, qe, nextO, nextD, excl, slot);
        if (wantShadow) {
            float4* e = b.shadow + 4u * qs;
            e[0] = make_float4(nextO.x, nextO.y, nextO.z, shD.x);
            e[1] = make_float4(shD.y, shD.z, dist, __uint_as_float(excl));
            e[2] = make_float4(Ld.x, Ld.y, Ld.z, __uint_as_float(slot));
            e[3] = make_float4(Tv.x, Tv.y, Tv.z, __uint_as_float(job));
        }
        if (wantEnd) ended[qd] = slot;
    }
    block_flush_rays(rc, counters);
}

// ------------------------------------------------------------- reorder --
// Counting-sort scatter of the next extension queue, after an exclusive scan
// turned the histogram into offsets; clears the histogram for the next
// iteration.  The order inside a bin varies from run to run; no path's
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

// ------------------------------------------------------------- resolve --
// Per covered pixel: the finished samples, in sample order, into the 8-bit
// running mean (incrementalAvg, Utils.cpp:48-71) and the FP32 sum; the first
// unfinished sample (stopped frame) ends the pixel's run.
__global__ __launch_bounds__(kBlock) void k_wf_resolve(RenderParams p, uint32_t spp, WfBuffers b,
                                                       uint32_t* __restrict__ bitmap, float* __restrict__ accum,
                                                       uint32_t* done_min) {
    const uint32_t ps = blockIdx.x * kBlock + threadIdx.x;
    if (ps >= b.pixel_slots) return;
    uint32_t x, y;
    slot_pixel(p, ps, &x, &y);
    if (!pixel_covered(p, x, y)) return;
    const uint32_t pid = y * p.width + x;
    const uint32_t bi = y * p.real_width + x;
    uint32_t bm = bitmap[bi];
    f3 acc{};
    if (accum) acc = f3{accum[3u * pid + 0u], accum[3u * pid + 1u], accum[3u * pid + 2u]};
    uint32_t s = b.first;
    for (; s < spp; ++s) {
        const float* r = b.rad + 3ull * ((s - b.first) * static_cast<uint64_t>(b.pixel_slots) + ps);
        if (__float_as_uint(r[0]) == kRadPending) break;
        const f3 c{r[0], r[1], r[2]};
        bm = incremental_avg(c, bm, s + 1u);
        acc = add(acc, c);
    }
    bitmap[bi] = bm;
    if (accum) {
        accum[3u * pid + 0u] = acc.x;
        accum[3u * pid + 1u] = acc.y;
        accum[3u * pid + 2u] = acc.z;
    }
    atomicMin(done_min, s);
}

// Folds the per-sample first hits of b.aov_s in sample order, as k_aov does
// over the same camera rays.
__global__ __launch_bounds__(kBlock) void k_wf_aov_resolve(RenderParams p, uint32_t spp, WfBuffers b,
                                                           float* __restrict__ aov) {
    const uint32_t ps = blockIdx.x * kBlock + threadIdx.x;
    if (ps >= b.pixel_slots) return;
    uint32_t x, y;
    slot_pixel(p, ps, &x, &y);
    if (!pixel_covered(p, x, y)) return;
    const uint32_t pid = y * p.width + x;
    f3 alb{}, nrm{};
    float dep = 0.0f;
    uint32_t s = b.first;
    for (; s < spp; ++s) {
        const float4* e = b.aov_s + 2ull * ((s - b.first) * static_cast<uint64_t>(b.pixel_slots) + ps);
        const float4 hv = e[1];
        if (__float_as_uint(hv.w) == kRadPending) break;
        if (hv.w == 0.0f) continue;
        const float4 av = e[0];
        alb = add(alb, f3{av.x, av.y, av.z});
        nrm = add(nrm, f3{hv.x, hv.y, hv.z});
        dep += av.w;
    }
    if (s == b.first) return;
    const float inv = 1.0f / static_cast<float>(s - b.first);
    float* o7 = aov + 7ull * pid;
    o7[0] = alb.x * inv;
    o7[1] = alb.y * inv;
    o7[2] = alb.z * inv;
    o7[3] = nrm.x * inv;
    o7[4] = nrm.y * inv;
    o7[5] = nrm.z * inv;
    o7[6] = dep * inv;
}

// ------------------------------------------------------------- launcher --
hipError_t launch_wavefront_aov_resolve(const RenderParams& p, uint32_t spp, const WfBuffers& b, float* aov,
                                        hipStream_t stream) {
    k_wf_aov_resolve<<<(b.pixel_slots + kBlock - 1u) / kBlock, kBlock, 0, stream>>>(p, spp, b, aov);
    return hipGetLastError();
}

hipError_t launch_wavefront_resolve(const RenderParams& p, uint32_t spp, const WfBuffers& b, uint32_t* bitmap,
                                    float* accum, uint32_t* done_min, hipStream_t stream) {
    k_wf_resolve<<<(b.pixel_slots + kBlock - 1u) / kBlock, kBlock, 0, stream>>>(p, spp, b, bitmap, accum, done_min);
    return hipGetLastError();
}

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

uint32_t wavefront_persistent_blocks(uint32_t num_cu, uint32_t cap, uint32_t groups, uint32_t share_pct) {
    static int perCu = 0;  // same for every device of this (single-arch) build
    if (perCu == 0) {
        int a = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, k_wf_trace_p<false, false, false>, kBlock, 0) != hipSuccess) {
            (void)hipGetLastError();
            return cap;
        }
        perCu = std::max(1, a);
    }
    const uint32_t resident = static_cast<uint32_t>(perCu) * num_cu;
    const uint32_t share = share_pct ? std::max(1u, resident * std::min(share_pct, 100u) / 100u)
                           : groups > 1u ? std::max(1u, resident * 6u / (5u * groups)) : resident;
    return std::min({cap, resident, share});
}

hipError_t launch_wavefront_start(const RenderParams& p, uint32_t first_sample, const WfBuffers& b, WfLoop* L,
                                  hipStream_t stream) {
    L->qin = b.queue_a;
    L->qout = b.queue_b;
    L->cin = kCntQueue0;
    L->cout = kCntQueue1;
    L->iter = 0;
    if (b.slot_count == 0u) return hipSuccess;
    k_wf_generate<<<(b.slot_count + kBlock - 1u) / kBlock, kBlock, 0, stream>>>(p, first_sample, b, L->qin,
                                                                                 b.counts + L->cin);
    return hipGetLastError();
}

hipError_t launch_wavefront_admit(const BatchDev& bd, const JobDev& jd, uint32_t job, uint32_t pool_wave,
                                  uint32_t job_wave, uint32_t n_waves, bool first, const WfBuffers& b,
                                  const WfLoop& L, hipStream_t stream) {
    if (n_waves == 0u) return hipSuccess;
    constexpr uint32_t wavesPerBlock = kBlock / 64u;
    k_wf_admit<<<(n_waves + wavesPerBlock - 1u) / wavesPerBlock, kBlock, 0, stream>>>(
        bd, jd, job, pool_wave, job_wave, n_waves, first ? 1u : 0u, b, L.qin, b.counts + L.cin);
    return hipGetLastError();
}

hipError_t launch_wavefront_iteration(const SceneDev& s, bool generic, bool persistent, bool count,
                                      const RenderParams& p, uint32_t spp, bool stop, uint32_t* bitmap, float* accum,
                                      const WfBuffers& b, WfLoop* L, uint32_t trace_blocks, uint32_t persist_blocks,
                                      const StackCfg& sc, unsigned long long* counters, uint32_t* err,
                                      hipEvent_t trace_begin, hipEvent_t trace_end, hipStream_t stream,
                                      const BatchDev* batch) {
    // A batch runs the persistent trace without instrumentation or reordering.
    if (batch && (!persistent || count || b.sort)) return hipErrorInvalidValue;
    const BatchDev bd = batch ? *batch : BatchDev{};
    uint32_t* c = b.counts;
    const uint32_t ePrev = (L->iter & 1u) ? kCntEnded0 : kCntEnded1;  // written by the previous shade
    const uint32_t eCur = (L->iter & 1u) ? kCntEnded1 : kCntEnded0;   // written by this iteration's shade
    uint32_t* endedPrev = b.ended[ePrev - kCntEnded0];
    uint32_t* endedCur = b.ended[eCur - kCntEnded0];
    if (trace_begin) {
        const hipError_t e = hipEventRecord(trace_begin, stream);
        if (e != hipSuccess) return e;
    }
    if (persistent && s.acc[MRT_CAT_TRIANGLE].kind == MRT_REGULAR_GRID) {
        if (batch) return hipErrorInvalidValue;
        if (count)
            k_wf_trace_grid_p<true><<<persist_blocks, kBlock, 0, stream>>>(
                s, L->qin, c + L->cin, c + kCntShadow, b.hits, b, c + L->cout, c + eCur, c + kCntFetch, sc, counters,
                err);
        else
            k_wf_trace_grid_p<false><<<persist_blocks, kBlock, 0, stream>>>(
                s, L->qin, c + L->cin, c + kCntShadow, b.hits, b, c + L->cout, c + eCur, c + kCntFetch, sc, counters,
                err);
    } else if (persistent) {
#define MRT_TRACE_P(C, M, O)                                                                                        \
    k_wf_trace_p<C, M, O><<<persist_blocks, kBlock, 0, stream>>>(s, L->qin, c + L->cin, c + kCntShadow, b.hits, b,  \
                                                                c + L->cout, c + eCur, c + kCntFetch, sc, counters, \
                                                                err, bd)
        if (batch) {
            if (b.overlap) MRT_TRACE_P(false, true, true); else MRT_TRACE_P(false, true, false);
        } else if (count) {
            if (b.overlap) MRT_TRACE_P(true, false, true); else MRT_TRACE_P(true, false, false);
        } else {
            if (b.overlap) MRT_TRACE_P(false, false, true); else MRT_TRACE_P(false, false, false);
        }
#undef MRT_TRACE_P
    } else {
#define MRT_EXTEND(G, C)                                                                                            \
    k_wf_extend<G, C><<<trace_blocks, kBlock, 0, stream>>>(s, L->qin, c + L->cin, b.hits, c + L->cout, c + eCur, sc, \
                                                           counters, err)
#define MRT_SHADOW(G, C) k_wf_shadow<G, C><<<trace_blocks, kBlock, 0, stream>>>(s, c + kCntShadow, b, sc, counters, err)
        if (generic) {
            if (count) { MRT_EXTEND(true, true); MRT_SHADOW(true, true); }
            else { MRT_EXTEND(true, false); MRT_SHADOW(true, false); }
        } else {
            if (count) { MRT_EXTEND(false, true); MRT_SHADOW(false, true); }
            else { MRT_EXTEND(false, false); MRT_SHADOW(false, false); }
        }
#undef MRT_EXTEND
#undef MRT_SHADOW
    }
    if (trace_end) {
        const hipError_t e = hipEventRecord(trace_end, stream);
        if (e != hipSuccess) return e;
    }
    // With reordering, regen and shade fill queue_b, which the scatter sorts
    // into queue_a; without it the two queues alternate.
    float4* fill = b.sort ? b.queue_b : L->qout;
    if (batch) {
        k_wf_regen<true><<<trace_blocks, kBlock, 0, stream>>>(p, spp, stop ? 1u : 0u, endedPrev, c + ePrev, b, fill,
                                                               c + L->cout, c + kCntShadow, c + kCntFetch,
                                                               c + kCntDone, bitmap, accum, bd);
        k_wf_shade<true><<<trace_blocks, kBlock, 0, stream>>>(s, p, L->qin, c + L->cin, b.hits, b, fill, c + L->cout,
                                                              c + kCntShadow, endedCur, c + eCur, counters, bd);
    } else {
        k_wf_regen<false><<<trace_blocks, kBlock, 0, stream>>>(p, spp, stop ? 1u : 0u, endedPrev, c + ePrev, b, fill,
                                                                c + L->cout, c + kCntShadow, c + kCntFetch,
                                                                c + kCntDone, bitmap, accum, bd);
        k_wf_shade<false><<<trace_blocks, kBlock, 0, stream>>>(s, p, L->qin, c + L->cin, b.hits, b, fill,
                                                               c + L->cout, c + kCntShadow, endedCur, c + eCur,
                                                               counters, bd);
    }
    if (b.sort) {
        size_t bytes = b.scan_bytes;
        const hipError_t e = hipcub::DeviceScan::ExclusiveSum(b.scan_temp, bytes, b.hist, b.offs, kSortBins, stream);
        if (e != hipSuccess) return e;
        k_wf_scatter<<<trace_blocks, kBlock, 0, stream>>>(c + L->cout, b);
    }
    wavefront_loop_advance(L, b.sort != 0u);
    return hipGetLastError();
}

void wavefront_loop_advance(WfLoop* L, bool sort) {
    const uint32_t eCur = (L->iter & 1u) ? kCntEnded1 : kCntEnded0;
    if (!sort) std::swap(L->qin, L->qout);
    std::swap(L->cin, L->cout);
    L->pending[0] = L->cin;
    L->pending[1] = kCntShadow;
    L->pending[2] = eCur;
    ++L->iter;
}

}  // namespace mrt
