// Host-callable launchers of the gfx950 kernels (mrt_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mrt_scene.hpp"

namespace mrt {

struct RenderParams;

// Global spill area of the traversal stacks: lane g of the launch owns
// entries spill[g + k * spill_stride], k < cap_spill.
struct StackCfg {
    uint32_t* spill;
    uint32_t spill_stride;
    uint32_t cap_spill;
};

hipError_t launch_trace_closest(const SceneDev& s, bool generic, const mrt_ray* rays, uint32_t n, mrt_hit* hits,
                                const StackCfg& sc, uint32_t cap_blocks, bool count, unsigned long long* counters,
                                uint32_t* err, hipStream_t stream);
hipError_t launch_trace_any(const SceneDev& s, bool generic, const mrt_ray* rays, uint32_t n, uint8_t* occ,
                            const StackCfg& sc, uint32_t cap_blocks, uint32_t* err, hipStream_t stream);
hipError_t launch_camera_rays(const mrt_camera& cam, const float* uvdd, uint32_t n, mrt_ray* out,
                              hipStream_t stream);
hipError_t launch_intersect_pairs(int cat, const void* prims, const mrt_ray* rays, uint32_t n, mrt_hit* closest,
                                  uint8_t* any, hipStream_t stream);
hipError_t launch_box_rays(const float* boxes6, const mrt_ray* rays, uint32_t n, uint8_t* hit, float* entry,
                           hipStream_t stream);
hipError_t launch_shade_rays(const SceneDev& s, bool generic, bool branches, const RenderParams& p,
                             const mrt_ray* rays, const int32_t* depths, uint32_t n, uint32_t sample, float* rgb,
                             uint8_t* flag, const StackCfg& sc, uint32_t cap_blocks, uint32_t* err,
                             hipStream_t stream);
uint32_t render_pass_blocks(const RenderParams& p);

// Scratch of the allocation-free trace entry point (mrt_trace_closest_noalloc):
// a file-scope device array of kNoAllocScratchWords words, one copy per GPU
// for the life of the process.  Words [0, kNoAllocReserved) hold the error
// flag, the rest the traversal-stack spill of one launch.
constexpr uint32_t kNoAllocScratchWords = 4u << 20;  // 16 MiB
constexpr uint32_t kNoAllocReserved = 64u;
// This GPU's address of the scratch array; the first call for a device (with
// that device current) looks it up, later ones return the kept address.
hipError_t noalloc_scratch(int device, uint32_t** out);
// Sets n words at p to v (a kernel rather than a memset node, so that the
// launch sequence can be captured).
hipError_t launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t stream);
hipError_t launch_aov(const SceneDev& s, bool generic, const RenderParams& p, uint32_t first, uint32_t spp, float* aov,
                      const StackCfg& sc, uint32_t* err, hipStream_t stream);
hipError_t launch_render_pass(const SceneDev& s, bool generic, bool branches, bool count, const RenderParams& p,
                              uint32_t sample, uint32_t* bitmap, float* accum, const StackCfg& sc,
                              unsigned long long* counters, uint32_t* err, hipStream_t stream);

}  // namespace mrt
