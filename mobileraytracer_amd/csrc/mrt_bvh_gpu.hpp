// GPU build of the reference BVH (the in-order SAH builder of BVH.hpp:64-155
// with getSplitIndex_SAH, BVH.hpp:274-311), node for node identical to the
// host restatement (mrt_build.cpp build_bvh_reference).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "mrt_build.hpp"

namespace mrt {

// Device scratch of a build, grown on demand and kept between builds.
struct BvhGpuScratch {
    void* p = nullptr;
    size_t bytes = 0;
    ~BvhGpuScratch() {
        if (p) (void)hipFree(p);
    }
};

// Builds over n boxes (6 floats each: min xyz, max xyz, device memory) with
// leaves of at most `leaf` (1..16) primitives.  Writes the reference node
// array (RefNode layout: box, offset, count) to `nodes` (device memory, at
// least 2n - 1 entries) and its length to *count.  Returns hipErrorNotSupported
// (nothing written) when a box holds a NaN coordinate or the levels built one
// by one exceed the limit for deep, unbalanced trees: the caller then builds
// on the host.  Synchronizes the stream.
hipError_t gpu_build_bvh(const float* boxes6, uint32_t n, uint32_t leaf, RefNode* nodes, uint32_t* count,
                         BvhGpuScratch* scratch, hipStream_t stream);

}  // namespace mrt
