// Library-internal host state shared by the C ABI sources: device buffers,
// error reporting, contexts and scenes.
#pragma once
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "mrt_build.hpp"
#include "mrt_bvh_gpu.hpp"
#include "mrt_grid_gpu.hpp"
#include "mrt_kernels.hpp"
#include "mrt_path.hpp"
#include "mrt_walk_gpu.hpp"
#include "mrt_wavefront.hpp"

namespace mrt {

// Message of the last failed call on this thread (mrt_last_error).
extern thread_local std::string g_error;

inline int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define MRT_HIP(expr)                                                                                        \
    do {                                                                                                     \
        const hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return ::mrt::fail(MRT_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device allocation owned by a scene, context or server.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int alloc(size_t n) {
        if (n <= bytes && p) return MRT_OK;
        release();
        if (n == 0) return MRT_OK;
        if (hipMalloc(&p, n) != hipSuccess) {
            p = nullptr;
            return fail(MRT_E_NOMEM, "hipMalloc of " + std::to_string(n) + " bytes failed");
        }
        bytes = n;
        return MRT_OK;
    }
    int upload(const void* src, size_t n, hipStream_t s) {
        const int rc = alloc(n);
        if (rc != MRT_OK) return rc;
        if (n) MRT_HIP(hipMemcpyAsync(p, src, n, hipMemcpyHostToDevice, s));
        return MRT_OK;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

constexpr uint32_t kStackLds = 16;      // entries per lane in LDS (mrt_kernels.hip)
constexpr uint32_t kLinearLights = 16;  // up to this many area lights are scanned linearly
constexpr uint32_t kGpuBuildMin = 65536; // MRT_BUILDER_AUTO builds triangle BVHs from this size on the GPU

// Render parameters and tile claim ranks of a validated render request
// (Renderer ctor tile sizes, Renderer.cpp:12-29; claim order,
// Renderer.cpp:79-88).
int render_setup(const mrt_scene* s, const mrt_render_desc* d, uint32_t stride_bytes, RenderParams* p,
                 std::vector<uint32_t>* rank);

// Whether the wavefront pipeline covers a request: the PathTracer with one
// light sample, no reflection/transmission, RayDepthMax >= 1, at most 65535
// samples (16-bit per-slot sample index) and fewer than 2^32 pixel samples.
bool wavefront_eligible(const mrt_scene* s, const mrt_render_desc* d);

// Stream of slot group g of a context: its stream, then its auxiliary streams
// (created on first use, default priority).  Callers hold ctx->mu.
int group_stream(mrt_ctx* c, uint32_t g, hipStream_t* out);

}  // namespace mrt

struct mrt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux[mrt::kWfMaxGroups - 1] = {};  // wavefront slot groups beyond the first
    uint32_t num_cu = 0;
    uint64_t budget = 0;
    std::atomic<uint32_t> children{0};  // live scenes and servers; the context outlives them
    std::mutex mu;
    // GPU BVH builds (mrt_bvh_build_gpu, scene builds): scratch, staged
    // boxes and the node array.
    mrt::BvhGpuScratch bvh_scratch;
    mrt::DevBuf bvh_boxes, bvh_nodes;
    mrt::WalkGpuScratch walk_scratch;  // device-side walk trees of GPU-built BVHs
    // GPU grid builds: scratch and the staged primitives.
    mrt::GridGpuScratch grid_scratch;
    mrt::DevBuf grid_prims;
    // This GPU's copy of the allocation-free entry point's scratch array.
    uint32_t* scratch = nullptr;
    mrt_tuning tuning{};  // zero fields: defaults
};

struct mrt_scene {
    mrt_ctx* ctx = nullptr;
    std::vector<mrt_triangle> tris;
    std::vector<mrt_sphere> sphs;
    std::vector<mrt_plane> plas;
    std::vector<mrt_material> mats;  // scene materials then one per light
    std::vector<mrt_light> lights;
    uint32_t n_scene_mats = 0;
    mrt::DevBuf tri_geo, tri_shade, sph, sph_mat, pla, mat, light_geo, light_shade, light_info;
    mrt::DevBuf nodes[4], perm[4], cell_off[3], cell_ref[3], walk_nodes, wide_nodes, cw_nodes, cw_recs;
    mrt::DevBuf spill, counters, err, stage_in, stage_out, stage_aux, tile_rank, wf_state, wf_rad, wf_aov, mega_ctl, mega_state,
        wf_group[mrt::kWfMaxGroups];
    // Host copy of the light tree (pairs relative to its own numbering) for
    // the combined walk tree; light slots follow the triangles in tri_geo.
    mrt::BvhOut light_bvh;
    uint32_t n_light_slots = 0;
    // Wavefront loop: pinned snapshots of the device counters (two, read
    // one batch behind the launches) and event pairs around trace launches.
    uint32_t* counts_host = nullptr;
    std::vector<hipEvent_t> trace_events;
    // Graph loop: the last captured batch of each slot group and the launch
    // arguments they were captured with (a frame with the same arguments
    // replays them).
    hipGraph_t wf_graph[mrt::kWfMaxGroups] = {};
    hipGraphExec_t wf_exec[mrt::kWfMaxGroups] = {};
    std::vector<uint8_t> wf_key;
    mrt::SceneDev dev{};
    ~mrt_scene() {
        for (uint32_t g = 0; g < mrt::kWfMaxGroups; ++g) {
            if (wf_exec[g]) (void)hipGraphExecDestroy(wf_exec[g]);
            if (wf_graph[g]) (void)hipGraphDestroy(wf_graph[g]);
        }
        if (counts_host) (void)hipHostFree(counts_host);
        for (hipEvent_t e : trace_events) (void)hipEventDestroy(e);
        if (counted && ctx) --ctx->children;
    }
    bool counted = false;  // holds one of ctx->children
    bool built = false;
    bool gpu_built = false;  // the triangle BVH came from the GPU build
    bool generic = false;
    uint32_t depth_need = 0;
    uint32_t servers = 0;  // render servers holding this scene (guarded by ctx->mu)
    std::atomic<bool> stop{false};
    std::atomic<uint32_t> progress{0};
};
