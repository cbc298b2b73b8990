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

struct WfBuffers {
    float4* sA;        // per slot: T.xyz, flags
    float4* sB;        // per slot: acc_all.xyz, pixel id
    float4* sC;        // per slot: acc_prefix.xyz
    float4* queue_a;   // 2 float4 per ray: o.xyz d.x | d.y d.z exclude slot  (traced)
    float4* queue_b;   // shade() output before reordering
    float4* hits;      // per queue entry: t, id, u, v
    float4* shadow;    // 4 float4 per shadow ray: o d | dist exclude | Ld slot | T
    uint32_t* keys;    // per queue_b entry: sort key
    uint32_t* hist;    // kSortBins counters (zero between bounces)
    uint32_t* offs;    // kSortBins exclusive offsets
    void* scan_temp;   // device-scan scratch
    size_t scan_bytes;
    uint32_t* counts;  // [0],[1] queue entries (ping-pong), [2] shadow entries
    float box_mn[3], box_inv[3];  // key quantisation box
    uint32_t sort;     // 1: reorder extension rays
};

// Slots (covered pixels rounded up to whole 8x8 wave blocks) of one pass.
uint32_t wavefront_slots(const RenderParams& p);

// Scratch bytes of the device scan over kSortBins counters.
size_t wavefront_scan_bytes();

// Bytes of the wavefront buffers for `slots` slots (scan scratch excluded).
inline size_t wavefront_bytes(uint32_t slots) {
    return static_cast<size_t>(slots) * sizeof(float4) * (3 + 2 + 2 + 1 + 4) +
           static_cast<size_t>(slots) * sizeof(uint32_t) + 2ull * kSortBins * sizeof(uint32_t) + 256;
}

// persistent: use the refilling extend / shadow kernels (BVH scenes of
// triangles and lights only).
hipError_t launch_wavefront_pass(const SceneDev& s, bool generic, bool persistent, bool count, const RenderParams& p,
                                 uint32_t sample, uint32_t* bitmap, float* accum, const WfBuffers& b,
                                 uint32_t trace_blocks, const StackCfg& sc, unsigned long long* counters,
                                 uint32_t* err, hipStream_t stream);

}  // namespace mrt
