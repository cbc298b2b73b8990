// GPU build of the RegularGrid cell lists, identical to the host build
// (mrt_build_grid.inl): same cells, same ascending primitive order per cell.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mobilert_gpu.h"
#include "mrt_math.hpp"

namespace mrt {

// Device scratch of grid builds, grown on demand and kept between builds.
struct GridGpuScratch {
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
    } counts, pairs;
    GridGpuScratch() = default;
    GridGpuScratch(const GridGpuScratch&) = delete;
    GridGpuScratch& operator=(const GridGpuScratch&) = delete;
    ~GridGpuScratch();
};

// Device arrays of a build (inside the scratch, valid until the next build):
// offsets (gs^3 + 1 entries) and refs (n_refs entries).
struct GridGpuOut {
    const uint32_t* offsets;
    const uint32_t* refs;
    uint32_t n_refs;
};

// Where a triangle's a, ab and ac start (in floats) in records of `stride`
// floats: mrt_triangle {28, 6, 3, 0}, packed (a, ab, ac) {9, 0, 3, 6}.
struct TriLayout {
    uint32_t stride, a, ab, ac;
};
constexpr TriLayout kTriAbi{28, 6, 3, 0}, kTriPacked{9, 0, 3, 6};

// Cell lists of n primitives of one category (device memory: triangles in
// layout L; spheres as mrt_sphere; planes as mrt_plane) on a grid of gs^3
// cells over bounds.  hipErrorNotSupported when a primitive box has a
// non-finite coordinate (the host build then decides).  Synchronizes the
// stream.
hipError_t gpu_build_grid(int category, const float* prims, TriLayout L, uint32_t n, const Box& bounds, int32_t gs,
                          GridGpuScratch* scratch, GridGpuOut* out, hipStream_t stream);

}  // namespace mrt
