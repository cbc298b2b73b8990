// Buffers and launcher of the wavefront path tracer (mrt_wavefront.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mrt_kernels.hpp"

namespace mrt {

struct RenderParams;

struct WfBuffers {
    float4* sA;        // per slot: T.xyz, flags
    float4* sB;        // per slot: acc_all.xyz, pixel id
    float4* sC;        // per slot: acc_prefix.xyz
    float4* queue_a;   // 2 float4 per ray: o.xyz d.x | d.y d.z exclude slot
    float4* queue_b;
    float4* hits;      // per queue entry: t, id, u, v
    float4* shadow;    // 4 float4 per shadow ray: o d | dist exclude | Ld slot | T
    uint32_t* counts;  // [0] queue_a entries, [1] queue_b entries, [2] shadow entries
};

// Slots (covered pixels rounded up to whole 8x8 wave blocks) of one pass.
uint32_t wavefront_slots(const RenderParams& p);

// Bytes of the wavefront buffers for `slots` slots.
inline size_t wavefront_bytes(uint32_t slots) {
    return static_cast<size_t>(slots) * sizeof(float4) * (3 + 2 + 2 + 1 + 4) + 64;
}

hipError_t launch_wavefront_pass(const SceneDev& s, bool generic, bool count, const RenderParams& p, uint32_t sample,
                                 uint32_t* bitmap, float* accum, const WfBuffers& b, uint32_t trace_blocks,
                                 const StackCfg& sc, unsigned long long* counters, uint32_t* err,
                                 hipStream_t stream);

}  // namespace mrt
