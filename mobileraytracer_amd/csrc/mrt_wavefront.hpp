// Buffers and launcher of the wavefront path tracer (mrt_wavefront.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mrt_kernels.hpp"
#include "mrt_path.hpp"

namespace mrt {

// Ray reordering between bounces: extension rays are counting-sorted by a
// key of 5-bit-per-axis Morton cell of the origin (inside the triangle BVH's
// root box) and the 3-bit direction octant.
constexpr uint32_t kSortBits = 18;
constexpr uint32_t kSortBins = 1u << kSortBits;

// Device counters (uint32 words of WfBuffers::counts).
enum : uint32_t {
    kCntQueue0 = 0,  // extension-queue entries (two queues, alternating)
    kCntQueue1 = 1,
    kCntShadow = 2,  // shadow-queue entries
    kCntEnded0 = 3,  // ended-slot lists (alternating)
    kCntEnded1 = 4,
    kCntDone = 5,    // finished samples (progress)
    kCntFetch = 8,   // fetch cursor of the persistent trace (own 32-byte sector)
    kCntWords = 16
};

// Buffers of one slot group.  The path state is indexed by the global slot;
// the queues, lists and counters belong to the group (slots [slot_base,
// slot_base + slot_count)), so groups run independently on their own streams.
struct WfBuffers {
    uint32_t slot_base, slot_count;
    float4* sA;        // per slot: T.xyz, flags
    float4* sB;        // per slot: acc_all.xyz, pixel id
    float4* sC;        // per slot: acc_prefix.xyz
    float4* queue_a;   // 2 float4 per ray: o.xyz d.x | d.y d.z exclude slot
    float4* queue_b;
    float4* hits;      // per traced extension entry: t, id, u, v
    float4* shadow;    // 4 float4 per shadow ray: o d | dist exclude | Ld slot | T
    uint32_t* ended[2];  // slots whose path ended, per shade (alternating)
    uint32_t* keys;    // per queue_b entry: sort key
    uint32_t* hist;    // kSortBins counters (zero between iterations)
    uint32_t* offs;    // kSortBins exclusive offsets
    void* scan_temp;   // device-scan scratch
    size_t scan_bytes;
    uint32_t* counts;  // kCntWords counters, zero before launch_wavefront_start
    float box_mn[3], box_inv[3];  // key quantisation box
    uint32_t sort;     // 1: reorder extension rays
    // Sample-parallel frames: `lanes` slots per pixel, slot k * pixel_slots +
    // ps running samples first + k, first + k + lanes, ... of pixel slot ps;
    // each finished sample's radiance goes to rad[(sample - first) *
    // pixel_slots + ps] (3 floats) and k_wf_resolve folds them into the
    // outputs in sample order.  lanes == 1 and rad == nullptr: samples are
    // folded as they finish.
    float* rad;
    uint32_t lanes, pixel_slots, first;
    // Frames with auxiliary outputs: each sample's first hit (albedo, t |
    // normal, hit flag) at aov_s[2 * ((sample - first) * pixel_slots + ps)],
    // written when its camera ray is shaded and folded by
    // launch_wavefront_aov_resolve; entries not written keep kRadPending in
    // the flag.  nullptr: none.
    float4* aov_s;
    // Persistent trace: a wave refills once refill_lanes lanes are idle and
    // reserves fetch_chunk (>= 64) queue entries per atomic.
    uint32_t refill_lanes, fetch_chunk;
    // Persistent trace: 1 issues each iteration's triangle and node loads
    // together (for launches too small to fill the chip).
    uint32_t overlap;
};

// Whether a persistent trace launch of `blocks` workgroups over at most
// `rays` rays runs the overlapped-load walk: when its rays cannot give every
// lane of the launch more than one ray (a ray's own walk latency then sets
// the launch time), unless `mode` (mrt_tuning::overlap_loads) forces it on
// (1) or off (2).
inline bool wavefront_overlap(uint64_t rays, uint32_t blocks, uint32_t mode) {
    if (mode == 1u) return true;
    if (mode == 2u) return false;
    return rays <= static_cast<uint64_t>(blocks) * 256u;
}

// Defaults of the launch-shape knobs (mrt_tuning).  Refill at 12 idle lanes:
// 1M soup, 1080p x 16 spp, 64-entry fetch runs, 8 / 10 / 12 / 14 / 16 / 24:
// 102.1 / 102.0 / 101.7 / 102.5 / 102.6 / 105.8 ms per frame.
constexpr uint32_t kDefaultRefillLanes = 12;
// 64 entries: one per lane of the wave's first refill, so a short queue is
// spread over as many waves as it has rays / 64 (256x256 x 1-sample frame
// 7.9 -> 6.7 ms against 128) and the 1080p x 16 frame is no slower (128 / 96 /
// 64: 103.3 / 103.1 / 102.5 ms).
constexpr uint32_t kDefaultFetchChunk = 64;
constexpr uint32_t kDefaultSlotLanes = 8;

// Radiance-buffer entries of samples not finished (x component bits; a NaN
// payload that arithmetic never produces).
constexpr uint32_t kRadPending = 0x7FBADBADu;

// A render job of a batch (the render server): the pool waves it owns carry
// its slots, and its scene's walk trees sit in the batch's shared arena.
struct JobDev {
    RenderParams p;
    uint32_t first_sample;  // samples [first_sample, spp) are rendered
    uint32_t spp;
    uint32_t stop;          // 1: start no new samples
    uint32_t scene;         // index into BatchDev::scenes
    uint32_t wide_root;     // arena wide node of the scene's closest-hit root (any-hit root: + 1)
    uint32_t walk_root;     // arena pair of the scene's closest-hit root (any-hit root: + 1)
    uint32_t* bitmap;       // RGBA8, p.real_width pixels per row
    float* accum;           // optional FP32 radiance sums
    // Sample-parallel jobs: job slot k * pixel_slots + ps runs samples
    // first_sample + k, + k + lanes, ... of pixel slot ps and leaves each
    // sample's radiance at rad[3 * ((sample - first_sample) * pixel_slots +
    // ps)] for the job's resolve.  rad == nullptr: lanes == 1, samples are
    // folded into bitmap / accum as they finish.
    float* rad;
    uint32_t lanes, pixel_slots;
};

// Device tables of a batch.  Every slot of pool wave w belongs to job
// waves[w].x and is slot (waves[w].y * 64 + lane) of that job's frame.
struct BatchDev {
    const JobDev* jobs;
    const SceneDev* scenes;
    uint2* waves;
    uint32_t* finished;  // per job: slots of this pipeline whose last sample is done
};

// Host side of the iteration loop.
struct WfLoop {
    float4* qin;          // extension queue traced next
    float4* qout;
    uint32_t cin, cout;   // their counters
    uint32_t iter;
    uint32_t pending[3];  // counters that are all zero once the frame is done
};

// Slots (covered pixels rounded up to whole 8x8 wave blocks).
uint32_t wavefront_slots(const RenderParams& p);

// Scratch bytes of the device scan over kSortBins counters.
size_t wavefront_scan_bytes();

// Bytes of the path state of `slots` slots (48 per slot).
inline size_t wavefront_state_bytes(uint32_t slots) { return static_cast<size_t>(slots) * sizeof(float4) * 3; }

// Bytes of one group's queues, lists and counters for `slots` slots (scan
// scratch excluded).
inline size_t wavefront_group_bytes(uint32_t slots) {
    return static_cast<size_t>(slots) * sizeof(float4) * (2 + 2 + 1 + 4) +
           static_cast<size_t>(slots) * sizeof(uint32_t) * 3 + 2ull * kSortBins * sizeof(uint32_t) + 512;
}

// Slot groups of a frame: independent pipelines on their own streams, so that
// one group's kernels fill the compute units while another's trace drains its
// slowest rays (a persistent trace launch ends only with its last ray).  On
// the 1M-triangle 1080p frame: 1 group 271 ms, 2 groups 219 ms, 3 groups
// 208 ms, 4 groups 299 ms (more streams than the process's 4 hardware queues
// serialise).
constexpr uint32_t kWfMaxGroups = 3;
inline uint32_t wavefront_groups(uint32_t slots, uint32_t max_groups = kWfMaxGroups) {
    return slots >= (1u << 16) ? std::max(1u, std::min(max_groups, kWfMaxGroups)) : 1u;
}

// Slots per pixel of a frame of `samples` samples over `pixel_slots` pixel
// slots: up to 8, while the frame stays within 2^24 slots (about 3.4 GB of
// path state and queues) and its per-sample radiance within 2 GiB.  More slots
// put more rays in each iteration, so the persistent lanes carry more rays per
// launch and a frame needs fewer iterations.  1M-triangle 1080p frame at 16
// spp: 1 / 2 / 4 / 8 / 16 slots per pixel took 187 / 169 / 152 / 140 / 137 ms
// (best launch share each).
inline uint32_t wavefront_lanes(uint32_t pixel_slots, uint32_t samples, uint32_t max_lanes = kDefaultSlotLanes) {
    uint32_t k = max_lanes;
    while (k > 1u && (k > samples || static_cast<uint64_t>(pixel_slots) * k > (1ull << 24) ||
                      12ull * samples * pixel_slots > (2ull << 30)))
        k /= 2u;
    return k;
}

// Workgroups of one persistent trace launch when `groups` launches run
// concurrently: 6 / (5 groups) of those that can be resident at once on
// `num_cu` compute units (all of them for one group), at most `cap`.  A launch
// carries a few rays per lane only if it does not claim the whole chip, and
// the regen and shade launches of the other groups need room beside it; on
// the 1M-triangle 1080p frame with three groups and one slot per pixel,
// launches of 25 / 34 / 40 / 46 / 67 / 100 % of the resident workgroups took
// 205 / 193 / 187 / 190 / 204 / 206 ms per frame; with 8 slots per pixel, the
// compressed tree and workgroup-aggregated queue appends, 30 / 40 / 50 / 60 /
// 100 % took 106.0 / 103.9 / 105.0 / 106.4 / 110.7 ms.
uint32_t wavefront_persistent_blocks(uint32_t num_cu, uint32_t cap, uint32_t groups, uint32_t share_pct = 0);

// Launch sizes of one slot group's iteration when at most `slots` of its
// slots hold paths: its queues hold at most one extension and one shadow ray
// per slot, so workgroups beyond ceil(2 * slots / 256) (trace) and
// ceil(slots / 256) (regen, shade) would find nothing to do.  A 256x256
// frame's iteration otherwise launches thousands of idle workgroups.
inline uint32_t sized_trace_blocks(uint32_t cap, uint64_t slots) {
    return static_cast<uint32_t>(std::max<uint64_t>(1u, std::min<uint64_t>(cap, (2u * slots + 255u) / 256u)));
}
inline uint32_t sized_blocks(uint32_t cap, uint64_t slots) {
    return static_cast<uint32_t>(std::max<uint64_t>(1u, std::min<uint64_t>(cap, (slots + 255u) / 256u)));
}

// Queues the camera rays of sample `first_sample` of every slot.
hipError_t launch_wavefront_start(const RenderParams& p, uint32_t first_sample, const WfBuffers& b, WfLoop* L,
                                  hipStream_t stream);

// Queues the camera rays of the first sample of job `job` (parameters `jd`,
// also stored as the job's table entry) on pool waves [pool_wave, pool_wave +
// n_waves), which become the job's waves [job_wave, job_wave + n_waves), into
// the queue the next iteration traces; `first` (the job's first run on this
// pipeline) also clears the pipeline's finished count of the job.
hipError_t launch_wavefront_admit(const BatchDev& bd, const JobDev& jd, uint32_t job, uint32_t pool_wave,
                                  uint32_t job_wave, uint32_t n_waves, bool first, const WfBuffers& b,
                                  const WfLoop& L, hipStream_t stream);

// Folds the per-sample radiance of a sample-parallel frame into bitmap and
// accumulation in sample order (Renderer.cpp:101-104 per sample), stopping at a
// pixel's first unfinished sample; counts[kCntDone] receives the smallest
// number of samples any covered pixel completed.
// Auxiliary outputs of a frame from b.aov_s: per covered pixel, the mean
// over samples [b.first, s) of the first-hit albedo, normal and distance
// (misses add nothing), s the first sample not written; 7 floats per pixel
// at aov + 7 * (y * width + x).
hipError_t launch_wavefront_aov_resolve(const RenderParams& p, uint32_t spp, const WfBuffers& b, float* aov,
                                        hipStream_t stream);
hipError_t launch_wavefront_resolve(const RenderParams& p, uint32_t spp, const WfBuffers& b, uint32_t* bitmap,
                                    float* accum, uint32_t* done_min, hipStream_t stream);

// One trace -> regen -> [reorder] -> shade iteration.  persistent: trace with
// the refilling kernel over the combined walk tree (BVH scenes of triangles
// and lights), `persist_blocks` workgroups; stop: finish ended paths without
// starting new samples.  trace_begin / trace_end (optional) bracket the trace
// launches.  With `batch`, the slots belong to the jobs of a batch (the job
// tables give each slot's parameters, scene and outputs; `s` holds the walk
// arena); p, spp, stop, bitmap and accum are then unused.
// Persistent megakernel (k_mega_p): one launch renders samples [first, spp)
// of every covered pixel; each lane takes work items (pixel slot ps, sample
// lane k: samples first + k, + k + lanes, ...) and keeps its path in
// registers.  lanes > 1: each sample's radiance goes to rad for
// launch_wavefront_resolve; lanes == 1: the lane folds its pixel's samples
// into bitmap / accum in order.  fetch and done are zero before the launch.
struct MegaArgs {
    RenderParams p;
    uint32_t first, spp;
    uint32_t lanes;
    uint32_t pixel_slots;
    float* rad;
    uint32_t* bitmap;
    float* accum;
    uint32_t* fetch;
    uint32_t* done;
    uint32_t service_lanes, refill_lanes, fetch_chunk;
    // Optional: 7 float4 per lane of the launch (SoA, stride = lanes) holding
    // each lane's path state between its walks instead of registers.
    float4* state;
};
hipError_t launch_mega_p(const SceneDev& s, bool count, const MegaArgs& a, uint32_t blocks, const StackCfg& sc,
                         unsigned long long* counters, uint32_t* err, hipStream_t stream);
// Workgroups of k_mega_p the chip holds at once.
uint32_t mega_blocks(uint32_t num_cu);
// Persistent megakernel defaults: shade once a quarter of the wave's lanes
// wait, refill at half, reserve 64 items per fetch (the fastest on the
// 1M-triangle soup from 256x256 x 1 to 1920x1080 x 16 samples).
constexpr uint32_t kDefaultServiceLanes = 16;
// AUTO renders frames of at least this many pixel samples with the wavefront.
constexpr uint64_t kMegaMaxPixelSamples = 6ull << 20;
constexpr uint32_t kDefaultMegaRefillLanes = 32;
constexpr uint32_t kDefaultMegaFetchChunk = 64;

// The host-side bookkeeping of one iteration (queue and counter roles,
// pending counters, iteration count), as launch_wavefront_iteration does it;
// for replays of captured iterations.
void wavefront_loop_advance(WfLoop* L, bool sort);

hipError_t launch_wavefront_iteration(const SceneDev& s, bool generic, bool persistent, bool count,
                                      const RenderParams& p, uint32_t spp, bool stop, uint32_t* bitmap, float* accum,
                                      const WfBuffers& b, WfLoop* L, uint32_t trace_blocks, uint32_t persist_blocks,
                                      const StackCfg& sc, unsigned long long* counters, uint32_t* err,
                                      hipEvent_t trace_begin, hipEvent_t trace_end, hipStream_t stream,
                                      const BatchDev* batch = nullptr);

}  // namespace mrt
