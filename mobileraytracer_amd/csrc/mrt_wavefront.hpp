// Buffers and launcher of the wavefront path tracer (mrt_wavefront.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mrt_kernels.hpp"

namespace mrt {

struct RenderParams;

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
inline uint32_t wavefront_groups(uint32_t slots) { return slots >= (1u << 16) ? kWfMaxGroups : 1u; }

// Workgroups of the persistent trace kernel: as many as can be resident at
// once on `num_cu` compute units, at most `cap`.
uint32_t wavefront_persistent_blocks(uint32_t num_cu, uint32_t cap);

// Queues the camera rays of sample `first_sample` of every slot.
hipError_t launch_wavefront_start(const RenderParams& p, uint32_t first_sample, const WfBuffers& b, WfLoop* L,
                                  hipStream_t stream);

// One trace -> regen -> [reorder] -> shade iteration.  persistent: trace with
// the refilling kernel over the combined walk tree (BVH scenes of triangles
// and lights), `persist_blocks` workgroups; stop: finish ended paths without
// starting new samples.  trace_begin / trace_end (optional) bracket the trace
// launches.
hipError_t launch_wavefront_iteration(const SceneDev& s, bool generic, bool persistent, bool count,
                                      const RenderParams& p, uint32_t spp, bool stop, uint32_t* bitmap, float* accum,
                                      const WfBuffers& b, WfLoop* L, uint32_t trace_blocks, uint32_t persist_blocks,
                                      const StackCfg& sc, unsigned long long* counters, uint32_t* err,
                                      hipEvent_t trace_begin, hipEvent_t trace_end, hipStream_t stream);

}  // namespace mrt
