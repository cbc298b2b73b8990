/*
 * mobilert_gpu.h - C ABI of the MI355X (gfx950) ray-scene library.
 *
 * Plain C linkage, plain pointers and sizes, opaque handles, int status codes
 * (0 = OK, < 0 = error; mrt_last_error() gives the message).  No exception
 * crosses this boundary.  Every compute entry point runs HIP kernels on the
 * device the context was created on; there is no CPU fallback: when no GPU is
 * present the context cannot be created and every call fails with
 * MRT_E_NODEVICE.
 *
 * Reference interfaces replaced (paths relative to the MobileRT source tree):
 *   Renderer::renderFrame / stopRender / getSample  app/MobileRT/Renderer.hpp:54-58
 *   Shader::rayTrace / shadowTrace                   app/MobileRT/Shader.hpp:73-75
 *   Shader::initializeAccelerators                   app/MobileRT/Shader.hpp:53
 *   BVH<T> / RegularGrid<T> / Naive<T> trace, shadowTrace
 *                                                    app/MobileRT/Accelerators/BVH.hpp:56-58,
 *                                                    RegularGrid.hpp:44-46, Naive.hpp:39-41, 55-74
 *   Triangle/Sphere/Plane::intersect                 app/MobileRT/Shapes/Triangle.hpp:55-57,
 *                                                    Sphere.hpp:35-37, Plane.hpp
 *   Camera::generateRay                              app/MobileRT/Camera.hpp:230-233
 *   extern "C" RayTrace                              app/System_dependent/Linux/c_wrapper.h:10-15
 *
 * Struct layouts mirror the reference classes field for field so that a C++
 * caller holding std::vector<Triangle> etc. can pass .data() directly.
 */
#ifndef MOBILERT_GPU_H
#define MOBILERT_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRT_API __attribute__((visibility("default")))

/* ---- status codes ------------------------------------------------------ */
enum mrt_status {
    MRT_OK = 0,
    MRT_E_INVALID = -1,     /* bad argument / size / handle                  */
    MRT_E_HIP = -2,         /* HIP runtime error                             */
    MRT_E_NOMEM = -3,       /* device allocation failed                      */
    MRT_E_UNSUPPORTED = -4, /* configuration not supported by the kernels    */
    MRT_E_BUDGET = -5,      /* admission refused: HBM budget exceeded        */
    MRT_E_STOPPED = -6,     /* render stopped by mrt_render_stop             */
    MRT_E_NODEVICE = -7,    /* no HIP device available                       */
    MRT_E_TIMEOUT = -8      /* mrt_server_wait: the job is not done yet       */
};

/* ---- enums that keep the reference's integer values -------------------- */
/* Shader::Accelerator, app/MobileRT/Shader.hpp:34-38 */
enum mrt_accelerator { MRT_NAIVE = 0, MRT_REGULAR_GRID = 1, MRT_BVH = 2 };
/* shader selector of the Linux CLI, app/System_dependent/Linux/c_wrapper.cpp:444-486 */
enum mrt_shader {
    MRT_SHADER_NOSHADOWS = 0,
    MRT_SHADER_WHITTED = 1,
    MRT_SHADER_PATHTRACER = 2,
    MRT_SHADER_DEPTHMAP = 3,
    MRT_SHADER_DIFFUSE = 4
};
/* Primitive categories in the order Shader::rayTrace visits them
 * (app/MobileRT/Shader.cpp:144-148 planes, spheres, triangles, then lights). */
enum mrt_category { MRT_CAT_PLANE = 0, MRT_CAT_SPHERE = 1, MRT_CAT_TRIANGLE = 2, MRT_CAT_LIGHT = 3 };
#define MRT_ID_NONE 0xFFFFFFFFu
#define MRT_MAKE_ID(cat, idx) ((((uint32_t)(cat)) << 30) | ((uint32_t)(idx)))
#define MRT_ID_CAT(id) ((uint32_t)(id) >> 30)
#define MRT_ID_INDEX(id) ((uint32_t)(id) & 0x3FFFFFFFu)

/* ---- scene data (reference class layouts) ------------------------------- */
/* Triangle stored form, app/MobileRT/Shapes/Triangle.hpp:16-25 (112 bytes). */
typedef struct mrt_triangle {
    float ac[3], ab[3], a[3];
    float na[3], nb[3], nc[3];
    float ca[3], cb[3], cc[3];
    int32_t material_id;
} mrt_triangle;

/* Sphere, app/MobileRT/Shapes/Sphere.hpp:16-18 (20 bytes). */
typedef struct mrt_sphere {
    float center[3];
    float sq_radius;
    int32_t material_id;
} mrt_sphere;

/* Plane, app/MobileRT/Shapes/Plane.hpp:16-18 (28 bytes). */
typedef struct mrt_plane {
    float normal[3];
    float point[3];
    int32_t material_id;
} mrt_plane;

/* Material, app/MobileRT/Material.hpp:13-17 (52 bytes). */
typedef struct mrt_material {
    float le[3], kd[3], ks[3], kt[3];
    float ior;
} mrt_material;

/* Sampler descriptor.  The reference samplers are host objects with atomic
 * counters or thread_local mt19937 streams (app/Components/Samplers,
 * app/MobileRT/Shader.cpp:193-236); on the GPU every draw is a pure function
 * of (seed, pixel, sample, dimension) so results do not depend on scheduling:
 *   CONSTANT          value (Components::Constant)
 *   RANDOM            counter-based hash stream (replaces MersenneTwister)
 *   HALTON            per-pixel scrambled radical inverse over the sample index
 *   STRATIFIED        per-pixel stratified over the sample index
 *   HALTON_COMPAT     the reference HaltonSeq counter: draw c of the frame, in
 *                     single-thread render order, is haltonSequence(c, 2) while
 *                     c < domain_size and 1.0 afterwards (HaltonSeq.cpp:14-22)
 *   STRATIFIED_COMPAT same with c / domain_size (Stratified.cpp:14-23)
 * The *_COMPAT kinds are accepted only for the pixel sampler. */
enum mrt_sampler_kind {
    MRT_SAMPLER_CONSTANT = 0,
    MRT_SAMPLER_RANDOM = 1,
    MRT_SAMPLER_HALTON = 2,
    MRT_SAMPLER_STRATIFIED = 3,
    MRT_SAMPLER_HALTON_COMPAT = 4,
    MRT_SAMPLER_STRATIFIED_COMPAT = 5
};
typedef struct mrt_sampler {
    int32_t kind;
    float value;          /* CONSTANT value                                  */
    uint32_t seed;        /* stream seed (RANDOM/HALTON/STRATIFIED)          */
    uint32_t domain_size; /* *_COMPAT: Sampler::domainSize_ (Sampler.cpp:9-13) */
} mrt_sampler;

/* Light: AreaLight (triangle emitter sampled with its own sampler,
 * app/Components/Lights/AreaLight.cpp:22-43) or PointLight (PointLight.cpp). */
enum mrt_light_kind { MRT_LIGHT_AREA = 0, MRT_LIGHT_POINT = 1 };
typedef struct mrt_light {
    int32_t kind;
    mrt_material radiance;  /* Light::radiance_, app/MobileRT/Light.hpp     */
    mrt_triangle triangle;  /* AreaLight::triangle_ (material_id -1)          */
    float position[3];      /* PointLight::position_                          */
    mrt_sampler sampler;    /* AreaLight::samplerPointLight_                  */
} mrt_light;

typedef struct mrt_scene_desc {
    const mrt_triangle* triangles; uint32_t n_triangles;
    const mrt_sphere* spheres;     uint32_t n_spheres;
    const mrt_plane* planes;       uint32_t n_planes;
    const mrt_material* materials; uint32_t n_materials;
    const mrt_light* lights;       uint32_t n_lights;
} mrt_scene_desc;

/* ---- rays and hits ------------------------------------------------------ */
/* Ray (app/MobileRT/Ray.hpp:14-17) in 32 bytes: the excluded primitive is an
 * id (category << 30 | index) instead of a pointer; tmax is the initial
 * Intersection::length_ (1e30 for rayTrace) or the shadowTrace distance. */
typedef struct mrt_ray {
    float org[3];
    float dir[3];
    float tmax;
    uint32_t exclude;
} mrt_ray;

/* Nearest hit: t, id (MRT_ID_NONE on a miss), barycentrics u (towards B) and
 * v (towards C) for triangles and area lights, 0 otherwise. */
typedef struct mrt_hit {
    float t;
    uint32_t id;
    float u;
    float v;
} mrt_hit;

/* ---- camera and render descriptors ------------------------------------- */
enum mrt_camera_kind { MRT_CAMERA_PERSPECTIVE = 0, MRT_CAMERA_ORTHOGRAPHIC = 1 };
/* Camera fields after construction (app/MobileRT/Camera.cpp:47-54):
 * perspective: fov_h/fov_v in radians (Perspective.cpp:12-18);
 * orthographic: half sizes sizeH_/sizeV_ (Orthographic.cpp:11-17). */
typedef struct mrt_camera {
    int32_t kind;
    float position[3], direction[3], right[3], up[3];
    float fov_h, fov_v;
} mrt_camera;

/* Kernel shape for mrt_render. */
/* MEGAKERNEL: every path on one lane.  Frames the wavefront kernels can run
 * (PathTracer, one light sample, no reflection or transmission, BVH) run the
 * persistent megakernel, one launch for the whole frame (desc.persistent 2:
 * one launch per sample pass instead); everything else runs one launch per
 * sample pass. */
enum mrt_kernel { MRT_KERNEL_AUTO = 0, MRT_KERNEL_MEGAKERNEL = 1, MRT_KERNEL_WAVEFRONT = 2 };

typedef struct mrt_render_desc {
    int32_t shader;            /* enum mrt_shader                              */
    uint32_t width, height;    /* Renderer width_/height_                      */
    uint32_t samples_pixel;    /* Renderer samplesPixel_                       */
    uint32_t samples_light;    /* Shader samplesLight_                         */
    int32_t ray_depth_min;     /* RayDepthMin (4 in Utils.hpp:35)              */
    int32_t ray_depth_max;     /* RayDepthMax (8 in Utils.hpp:36)              */
    mrt_camera camera;
    mrt_sampler sampler_pixel; /* Renderer samplerPixel_                       */
    mrt_sampler sampler_rr;    /* PathTracer samplerRussianRoulette_           */
    mrt_sampler sampler_light; /* Shader::getLight draw                        */
    mrt_sampler sampler_hemi;  /* Shader::getCosineSampleHemisphere draws      */
    float tile_values[256];    /* Camera VALUES table in claim order           */
    int32_t kernel;            /* enum mrt_kernel                              */
    uint32_t first_sample;     /* first sample pass (progressive resume)       */
    uint32_t instrument;       /* 1: also count node visits / primitive tests  */
    uint32_t ray_sort;         /* wavefront: 0 default (off), 1 on, 2 off:
                                  extension rays in origin-cell order          */
    uint32_t persistent;       /* wavefront: 0 default (on), 1 on, 2 off: lanes
                                  refill from the queue (BVH, triangles+lights) */
    float depth_max_point[3];  /* DepthMap maxPoint_ (DepthMap.cpp:12-16)       */
    uint32_t tile_skip[8];     /* bit t (t = tile row * 16 + tile column) set:
                                  tile t is not rendered (region renders and
                                  tile shards; no counterpart in the
                                  reference); all zero: every tile.  Other
                                  pixels are unchanged by it.                 */
} mrt_render_desc;

typedef struct mrt_render_stats {
    uint64_t rays_camera;      /* primary rays traced                          */
    uint64_t rays_extension;   /* diffuse continuation rays traced             */
    uint64_t rays_shadow;      /* shadowTrace calls                            */
    uint64_t rays_specular;    /* reflection + transmission rays traced        */
    uint64_t node_visits;      /* interior node pairs fetched (instrumented)   */
    uint64_t prim_tests;       /* primitive tests (instrumented)               */
    uint64_t light_node_visits, light_prim_tests;    /* of which: light BVH / scan  */
    uint64_t shadow_node_visits, shadow_prim_tests;  /* of which: shadowTrace       */
    uint32_t samples_done;     /* completed samples (every covered pixel)      */
    uint32_t covered_pixels;   /* pixels inside the 16x16 tile grid            */
    float ms_render;           /* device time of the render, milliseconds      */
    float ms_kernel;           /* device time of the sample loop (megakernel:
                                  summed pass launches; wavefront: the whole
                                  iteration loop)                              */
    float ms_trace;            /* wavefront: summed trace launches             */
    uint32_t trace_launches;   /* wavefront: trace launches (iterations)       */
    uint64_t trace_lane_iters; /* persistent trace, instrumented: iterations in
                                  which a lane held a ray ...                  */
    uint64_t trace_lane_slots; /* ... and all lane iterations (SIMD use)       */
    float ms_trace_busy;       /* wavefront: time at least one trace launch ran
                                  (launches of slot groups overlap)            */
    uint32_t trace_groups;     /* wavefront: slot groups run concurrently      */
    uint32_t slot_lanes;       /* wavefront, persistent megakernel: path slots
                                  (sample lanes) per pixel                     */
    uint32_t persistent_megakernel; /* 1: the frame ran as one persistent
                                  megakernel launch (ms_kernel is its time)    */
} mrt_render_stats;

/* Accelerator build parameters (Shader::initializeAccelerators inputs). */
typedef struct mrt_accel_params {
    float camera_min[3];       /* Camera::getAABB(), joins the grid bounds     */
    float camera_max[3];       /* (app/MobileRT/Shader.cpp:52-62)              */
    uint32_t leaf_size;        /* BVH maxLeafSize (2 in BVH.hpp:19)            */
    uint32_t builder;          /* mrt_builder: where the BVH is built          */
    uint32_t reserved[2];
} mrt_accel_params;
/* Where mrt_scene_build builds.  BVH, HOST: the reference builder's
 * restatement and a binned-SAH 4-wide walk tree over its leaves; GPU: the
 * same binary tree built on the GPU (mrt_bvh_build_gpu) and a 4-wide walk
 * tree collapsed from its pairs (each wide node = a pair with its interior
 * children expanded), all written on the device (the host adds the roots
 * and the light part); both walk trees give the same hits (every interior box
 * is the exact union of the reference leaves below it).  RegularGrid, GPU:
 * the cell lists of mrt_grid_build_gpu, identical to the host's.  AUTO: GPU
 * for categories of 65536 primitives or more, HOST below.  A GPU build that
 * declines (NaN / non-finite coordinates, more than 1024 BVH levels above 64
 * primitives) falls back to HOST. */
enum mrt_builder { MRT_BUILDER_AUTO = 0, MRT_BUILDER_HOST = 1, MRT_BUILDER_GPU = 2 };

typedef struct mrt_ctx mrt_ctx;
typedef struct mrt_scene mrt_scene;

/* Memory location of the ray/hit/bitmap/accumulation arguments. */
enum mrt_mem { MRT_MEM_HOST = 0, MRT_MEM_DEVICE = 1 };

/* ---- context ------------------------------------------------------------ */
MRT_API const char* mrt_last_error(void);
MRT_API int mrt_device_count(int* count);
MRT_API int mrt_ctx_create(int device, mrt_ctx** out);
/* Fails with MRT_E_INVALID while scenes or render servers of the context live. */
MRT_API int mrt_ctx_destroy(mrt_ctx* ctx);
/* HBM budget for scene admission, bytes (0 = unlimited). Mirrors the memory
 * admission of app/System_dependent/Android_JNI/DrawView.cpp:373-416. */
MRT_API int mrt_ctx_set_budget(mrt_ctx* ctx, uint64_t bytes);
MRT_API int mrt_ctx_synchronize(mrt_ctx* ctx);
/* Launch-shape knobs of the wavefront renderer and the render servers a
 * context creates afterwards (no counterpart in the reference; they change
 * speed only, never a pixel).  Zero fields keep the defaults. */
typedef struct mrt_tuning {
    uint32_t refill_lanes;     /* persistent trace: a wave takes new rays once this
                                  many of its lanes are idle (1-64; 12, the
                                  persistent megakernel 32)                    */
    uint32_t fetch_chunk;      /* queue entries a wave reserves per atomic
                                  (64-4096; 64)                                */
    uint32_t launch_share_pct; /* workgroups of one slot group's trace launch, %
                                  of those the chip holds at once (1-100; 100
                                  for one group, 120 / groups for more)        */
    uint32_t max_slot_lanes;   /* path slots per pixel at most (1, 2, 4, 8, 16; 8) */
    uint32_t max_groups;       /* slot groups of a frame at most (1-3; 3)      */
    uint32_t overlap_loads;    /* persistent trace issues each step's triangle and
                                  node loads together: 0 when a launch cannot
                                  give each lane more than one ray, 1 always,
                                  2 never                                      */
    uint32_t graph_loop;       /* 1: mrt_render's wavefront loop as captured
                                  HIP graphs, one per slot group (4 of its
                                  iterations, replayed on the group's stream
                                  until the group drains); the scene keeps
                                  them, and a frame with the same launch
                                  arguments replays them without a new
                                  capture (graph frames report no per-launch
                                  trace times).  0, 2: launched iterations    */
    uint32_t service_lanes;    /* persistent megakernel: a wave shades once this
                                  many of its lanes finished their walks (1-64;
                                  16)                                          */
} mrt_tuning;
MRT_API int mrt_ctx_set_tuning(mrt_ctx* ctx, const mrt_tuning* tuning);
MRT_API int mrt_ctx_get_tuning(const mrt_ctx* ctx, mrt_tuning* tuning);

/* ---- scene -------------------------------------------------------------- */
/* Copies the scene to the device (Shader(Scene&&, ...), Shader.cpp:18-25). */
MRT_API int mrt_scene_create(mrt_ctx* ctx, const mrt_scene_desc* desc, mrt_scene** out);
/* Fails with MRT_E_INVALID while a render server holds the scene. */
MRT_API int mrt_scene_destroy(mrt_scene* scene);
/* Builds the accelerator of every category (Shader::initializeAccelerators,
 * Shader.cpp:31-122): BVH with the reference in-order SAH builder
 * (BVH.hpp:64-155, 274-311), RegularGrid with the reference sizing and cell
 * registration (Shader.cpp:40-113, RegularGrid.hpp:50-154), or Naive. */
MRT_API int mrt_scene_build(mrt_scene* scene, int accelerator, const mrt_accel_params* params);
/* Device bytes held by the scene (geometry + accelerator). */
MRT_API int mrt_scene_footprint(const mrt_scene* scene, uint64_t* bytes);

/* ---- traversal entry points -------------------------------------------- */
/* Shader::rayTrace's intersection step (Shader.cpp:124-160): nearest hit over
 * planes, spheres, triangles (through the built accelerator) and lights
 * (Scene::traceLights, Scene.cpp:31-37). */
MRT_API int mrt_trace_closest(mrt_scene* scene, const mrt_ray* rays, uint32_t n, mrt_hit* hits, int mem);
/* mrt_trace_closest on device memory that allocates nothing.  No counterpart
 * in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops).  rays and hits are
 * 16-byte-aligned device pointers.  The traversal-stack spill and the error
 * flag live in a 16 MiB array the library holds on each GPU until the process
 * ends, so the call allocates nothing; trees deeper than it supports (about
 * 16,000 levels) are refused with MRT_E_BUDGET.  Results equal
 * mrt_trace_closest's.  Calls on one GPU share that array: calls that overlap
 * on one GPU (scenes of different contexts on it, used from several threads)
 * race. */
MRT_API int mrt_trace_closest_noalloc(mrt_scene* scene, const mrt_ray* rays, uint32_t n, mrt_hit* hits);
/* Shader::shadowTrace (Shader.cpp:162-187): any hit in (1e-6, tmax), lights excluded. */
MRT_API int mrt_trace_any(mrt_scene* scene, const mrt_ray* rays, uint32_t n, uint8_t* occluded, int mem);
/* Shader::rayTrace including shade(): radiance and the "hit light" flag for a
 * batch of rays, with the render descriptor's shader, depths and samplers;
 * ray i uses pixel index i and sample `sample` for its sampler draws. */
MRT_API int mrt_shade_rays(mrt_scene* scene, const mrt_render_desc* desc, const mrt_ray* rays,
                           const int32_t* depths, uint32_t n, uint32_t sample,
                           float* rgb, uint8_t* hit_light, int mem);
/* Camera::generateRay for a batch of (u, v, deviationU, deviationV). */
MRT_API int mrt_camera_rays(mrt_ctx* ctx, const mrt_camera* camera, const float* uvdd, uint32_t n,
                            mrt_ray* rays, int mem);
/* Pairwise primitive tests: ray i against primitive i (KAT entry points for
 * Triangle::intersect (both forms) and Plane::intersect, Sphere::intersect). */
MRT_API int mrt_intersect_pairs(mrt_ctx* ctx, int category, const void* prims, const mrt_ray* rays,
                                uint32_t n, mrt_hit* closest, uint8_t* any, int mem);
/* ::MobileRT::intersect(AABB, Ray) (AABB.cpp:15-35): box i (min xyz, max xyz)
 * against ray i; hit[i] = 1 when the slab test passes, entry[i] (optional)
 * = max(tmin, 0). */
MRT_API int mrt_box_rays(mrt_ctx* ctx, const float* boxes6, const mrt_ray* rays, uint32_t n, uint8_t* hit,
                         float* entry, int mem);

/* ---- render ------------------------------------------------------------- */
/* Renderer::renderFrame (Renderer.cpp:31-59, 67-116): renders desc->samples_pixel
 * passes over the 16x16 tile grid, folding every pass into the RGBA8 bitmap
 * with the reference 8-bit running mean (Utils.cpp:48-71), and, when accum is
 * not NULL, adding the FP32 radiance of every pass to accum (3 floats per
 * pixel, stride width).  stride_bytes is the bitmap row stride. */
MRT_API int mrt_render(mrt_scene* scene, const mrt_render_desc* desc, uint32_t* bitmap,
                       uint32_t stride_bytes, float* accum, int mem, mrt_render_stats* stats);
/* Renderer::stopRender (Renderer.cpp:61-65): callable from any thread; the
 * running mrt_render starts no further sample (megakernel: pass; wavefront:
 * per-pixel sample), lets the ones in flight finish and returns
 * MRT_E_STOPPED unless every sample was already done. */
MRT_API int mrt_render_stop(mrt_scene* scene);
/* Renderer::getSample (Renderer.cpp:118-120). */
MRT_API int mrt_render_progress(const mrt_scene* scene, uint32_t* samples_done);
/* Auxiliary outputs of the frame desc describes, for a denoiser (no
 * counterpart in the reference): per covered pixel, over the camera rays of
 * samples [first_sample, samples_pixel) (the render's jitter), the mean
 * first-hit albedo (the material's Kd, the vertex colour without a material,
 * Le on an emitter), shading normal and hit distance (zeros for a miss).
 * aov: 7 floats per pixel (albedo.xyz, normal.xyz, depth), width x height,
 * pixels outside the tile grid zero. */
MRT_API int mrt_render_aovs(mrt_scene* scene, const mrt_render_desc* desc, float* aov, int mem);
/* mrt_render plus the outputs of mrt_render_aovs for the same desc, in one
 * call (no counterpart in the reference): the wavefront pipeline records each
 * sample's first hit when it shades the camera ray (per-sample records, up
 * to 2 GiB of them) and folds them in sample order after the frame, so aov
 * is identical to mrt_render_aovs's without tracing the camera rays twice;
 * the megakernel, or frames beyond that size, run the AOV kernel after the
 * render.  A stopped render's aov covers the samples finished in every
 * pixel. */
MRT_API int mrt_render_with_aovs(mrt_scene* scene, const mrt_render_desc* desc, uint32_t* bitmap,
                                 uint32_t stride_bytes, float* accum, float* aov, int mem, mrt_render_stats* stats);

/* ---- render server (continuous batching) -------------------------------- */
/* No counterpart in the reference, whose Renderer renders one frame at a time
 * (Renderer.cpp:31-59); this is the serving loop around it.  A server keeps
 * scenes resident (their walk trees in one device arena) and runs the render
 * jobs submitted to it in shared wavefront pipelines: a job is admitted, in
 * submission order, as soon as enough path slots and job memory are free; its
 * paths then run in the same kernel launches as those of the jobs already
 * running, and its bitmap / accumulation end up exactly as mrt_render leaves
 * them for the same request.  A job admitted while at least twice its
 * pixel slots are free runs up to 8 slots per pixel (its samples in
 * parallel, each sample's radiance kept in the sample memory and folded in
 * sample order when the job finishes).  Jobs need the wavefront pipeline (PathTracer,
 * samples_light 1, no reflection/transmission, scenes of triangles and lights
 * with a BVH).  All calls are thread-safe. */
typedef struct mrt_server mrt_server;

typedef struct mrt_server_desc {
    uint32_t max_slots;        /* path slots (0: 33554432, sixteen 1920x1080
                                  frames)                                      */
    uint32_t max_jobs;         /* jobs admitted at once (0: 256)               */
    uint32_t max_scenes;       /* resident scenes (0: 64)                      */
    uint32_t pipelines;        /* concurrent pipelines, 1..3 (0: 3)            */
    uint64_t job_memory;       /* device bytes for the outputs of admitted
                                  jobs (host-memory jobs stage their bitmap and
                                  accumulation there) (0: 1 GiB)               */
    uint64_t scene_memory;     /* device bytes for the scene arena (0: no limit
                                  beyond the context budget)                   */
    uint64_t sample_memory;    /* device bytes for the per-sample radiance of
                                  sample-parallel jobs (0: 8 GiB)              */
    uint32_t idle_direct;      /* 1: a job that reaches an idle server (no job
                                  admitted or queued beside it) and that
                                  mrt_render's AUTO kernel would give the
                                  one-launch persistent megakernel (below
                                  6 x 2^20 pixel samples) renders alone
                                  through mrt_render on the server's worker,
                                  with the same results; jobs submitted
                                  meanwhile wait for it (a cancel lets it
                                  complete, then it fails with
                                  MRT_E_STOPPED).  0 (default): every job runs
                                  on the pipelines                             */
    uint32_t reserved;
} mrt_server_desc;

enum mrt_job_state { MRT_JOB_QUEUED = 0, MRT_JOB_RUNNING = 1, MRT_JOB_DONE = 2, MRT_JOB_FAILED = 3 };

typedef struct mrt_job_info {
    int32_t state;             /* enum mrt_job_state                           */
    int32_t status;            /* MRT_OK, MRT_E_STOPPED after mrt_server_cancel,
                                  or the error of a failed job                 */
    double t_submit_us;        /* host steady clock, microseconds: submitted,  */
    double t_admit_us;         /* admitted (0 while queued),                   */
    double t_gpu_done_us;      /* last sample seen done on the device,         */
    double t_done_us;          /* result in the caller's buffers               */
    uint32_t lanes;            /* slots per pixel the job ran with (1 while
                                  queued; > 1: sample-parallel)               */
    uint32_t direct;           /* 1: rendered alone by an idle server
                                  (mrt_server_desc.idle_direct)               */
} mrt_job_info;

typedef struct mrt_serving_stats {
    uint64_t jobs_submitted, jobs_done, jobs_failed;
    uint64_t rays_camera, rays_extension, rays_shadow;  /* all jobs so far      */
    uint32_t jobs_queued, jobs_running;
    uint32_t slots_total, slots_used;
    uint64_t scene_bytes;      /* scene arena in use                           */
    uint64_t job_bytes_used;   /* job memory in use                            */
    uint64_t batches;          /* iteration batches issued                     */
} mrt_serving_stats;

MRT_API int mrt_server_create(mrt_ctx* ctx, const mrt_server_desc* desc, mrt_server** out);
/* Cancels queued jobs, lets running ones finish, then frees the server. */
MRT_API int mrt_server_destroy(mrt_server* server);
/* Makes a built scene resident (copies its walk trees into the arena).  The
 * scene must outlive the server; mrt_scene_destroy refuses it meanwhile. */
MRT_API int mrt_server_add_scene(mrt_server* server, mrt_scene* scene, uint32_t* scene_id);
/* Queues a render of a resident scene: the arguments mean what they mean for
 * mrt_render.  Host buffers are read at admission and written when the job
 * is done; they must stay valid until then. */
MRT_API int mrt_server_submit(mrt_server* server, uint32_t scene_id, const mrt_render_desc* desc,
                              uint32_t* bitmap, uint32_t stride_bytes, float* accum, int mem,
                              uint64_t* job_id);
/* Waits until the job is done or failed (timeout_us < 0: no limit) and
 * returns its status (MRT_E_TIMEOUT on time-out); the job is forgotten once a
 * wait has returned its end. */
MRT_API int mrt_server_wait(mrt_server* server, uint64_t job_id, int64_t timeout_us, mrt_job_info* info);
MRT_API int mrt_server_poll(mrt_server* server, uint64_t job_id, mrt_job_info* info);
/* Renderer::stopRender for one job: queued, it fails with MRT_E_STOPPED;
 * running, its pixels finish the sample in flight, the samples done so far
 * reach the caller's buffers and it fails with MRT_E_STOPPED. */
MRT_API int mrt_server_cancel(mrt_server* server, uint64_t job_id);
MRT_API int mrt_server_stats(mrt_server* server, mrt_serving_stats* stats);

/* ---- learned denoiser ------------------------------------------------------ */
/* No counterpart in the reference (it writes the 8-bit running mean of the
 * samples, Utils.cpp:48-71, and stops there).  A U-Net over a low-sample
 * frame: radiance plus the mrt_render_aovs outputs (10 channels) in, radiance
 * out.  BF16 weights and activations, 3x3 convolutions on MFMA with FP32
 * accumulation.
 *
 * Network, for L levels of c[0..L-1] channels (multiples of 32, at most 256):
 *   input  x0 = [log(1 + max(r * radiance_scale, 0)) (3), albedo (3),
 *                normal (3), depth * depth_scale (1)], frame zero-padded
 *                right and bottom to multiples of 2^(L-1);
 *   encoder level l: e_l = relu(conv(relu(conv(in_l)))) with c[l] channels,
 *                in_0 = x0, in_l = maxpool2x2(e_{l-1});
 *   decoder level l = L-2 .. 0: d_l = relu(conv(relu(conv(
 *                concat(upsample2x_nearest(d_{l+1}), e_l))))) with c[l]
 *                channels, d_{L-1} = e_{L-1};
 *   output y = conv(d_0) (3 channels, no activation); radiance
 *                max(exp(y) - 1, 0), cropped to the frame.
 * Weights: FP32, layer after layer in the order encoder 0..L-1 (two convs
 * each), decoder L-2..0 (two each), output conv: each layer its weight
 * [cout][cin][3][3] then its bias [cout] (input channels of a decoder's
 * first conv: the upsampled channels, then the skip channels).  The library
 * rounds them to BF16 (nearest even). */
#define MRT_DENOISER_MAX_LEVELS 5
typedef struct mrt_denoiser mrt_denoiser;

typedef struct mrt_denoiser_desc {
    uint32_t levels;                            /* 1..5 (0: 5)                       */
    uint32_t channels[MRT_DENOISER_MAX_LEVELS]; /* per level (all 0: 32 64 96 128 128) */
    float depth_scale;                          /* multiplies the depth AOV (0: 1)   */
    uint32_t reserved[5];
} mrt_denoiser_desc;

/* Number of FP32 weights (with biases) of the network desc describes. */
MRT_API int mrt_denoiser_weight_count(const mrt_denoiser_desc* desc, uint64_t* count);
MRT_API int mrt_denoiser_create(mrt_ctx* ctx, const mrt_denoiser_desc* desc, const float* weights, uint64_t count,
                                mrt_denoiser** out);
/* Denoises a width x height frame: radiance (3 floats per pixel, multiplied
 * by radiance_scale, e.g. 1/spp for an mrt_render accumulation) and aov (7
 * floats per pixel, as mrt_render_aovs writes) in, out (3 floats per pixel)
 * out; all three in host or device memory per mem.  Activations stay on the
 * device between calls and grow with the largest frame seen. */
MRT_API int mrt_denoiser_run(mrt_denoiser* dn, const float* radiance, float radiance_scale, const float* aov,
                             uint32_t width, uint32_t height, float* out, int mem);
/* Milliseconds of the last run's device work (convolutions and the
 * packing kernels), from events around it. */
MRT_API int mrt_denoiser_last_ms(const mrt_denoiser* dn, float* ms);
MRT_API int mrt_denoiser_destroy(mrt_denoiser* dn);
/* One layer of the network on its own: a 3x3 convolution (zero padding 1) of
 * x (height x width x cin BF16, NHWC) with w ([3][3][cout][cin] BF16, tap
 * (dy, dx) major) and bias (cout FP32, or NULL), ReLU when relu is set, into
 * y (height x width x cout BF16).  cin and cout multiples of 32.  Device
 * buffers (mem = MRT_MEM_DEVICE): x and w 16-byte aligned, y 8-byte aligned,
 * bias 4-byte aligned (MRT_E_INVALID otherwise).  Each call reorders w into
 * the kernel's fragment order in a device buffer of the same size that it
 * allocates and frees. */
MRT_API int mrt_conv3x3_bf16(mrt_ctx* ctx, const uint16_t* x, uint32_t height, uint32_t width, uint32_t cin,
                             const uint16_t* w, const float* bias, uint32_t cout, int relu, uint16_t* y, int mem);

/* ---- denoiser training ----------------------------------------------------- */
/* No counterpart in the reference.  Trains the network above on pairs the
 * renderer produced: input = the packed network input of a low-sample frame
 * (10 FP32 per pixel: log(1 + max(radiance, 0)) (3), albedo (3), normal (3),
 * depth * depth_scale (1)), target = log(1 + max(radiance, 0)) of an
 * independent high-sample frame (3 FP32 per pixel).  A step takes `batch`
 * square crops of side `crop` (frame, position and left-right / top-bottom
 * mirroring drawn from the seed and the step index), runs the network in
 * BF16 (weights and activations rounded to BF16 as mrt_denoiser_run does,
 * the output y rounded to BF16 too), the loss mean((y - target)^2) over the
 * batch's pixels and 3 channels, the gradients (BF16 activation gradients,
 * FP32 accumulation, FP32 weight gradients; weights carried straight through
 * their BF16 rounding) and one Adam update of FP32 master weights
 * (torch.optim.Adam's arithmetic with its double-precision hyperparameters
 * rounded where torch rounds them, no weight decay).  Kernels: BF16 MFMA for
 * every convolution's forward, backward-data and weight-gradient products. */
typedef struct mrt_denoiser_trainer mrt_denoiser_trainer;

typedef struct mrt_denoiser_train_desc {
    uint32_t batch;   /* crops per step, 1..64 (0: 8)                                    */
    uint32_t crop;    /* crop side, a multiple of 8 * 2^(levels - 1), at most 4096 (0: 256) */
    double beta1;     /* Adam (0: 0.9)                                                   */
    double beta2;     /* (0: 0.999)                                                      */
    double eps;       /* (0: 1e-8)                                                       */
    uint64_t seed;    /* crop choice                                                     */
    uint32_t reserved[8];
} mrt_denoiser_train_desc;

typedef struct mrt_denoiser_train_stats {
    uint64_t steps;        /* Adam steps taken                                         */
    double last_loss;      /* mean loss of the last mrt_denoiser_trainer_step call     */
    float last_ms;         /* device milliseconds of that call (events around it)      */
    float ms_per_step;     /* last_ms / its steps                                      */
    double flops_per_step; /* algorithmic FLOP of one step: forward, backward-data
                              (all layers but the first) and weight-gradient products,
                              2 * 9 * cin * cout per pixel, logical channels           */
    uint32_t reserved[8];
} mrt_denoiser_train_stats;

/* weights: the initial FP32 weights in mrt_denoiser_create's layout. */
MRT_API int mrt_denoiser_trainer_create(mrt_ctx* ctx, const mrt_denoiser_desc* net, const float* weights,
                                        uint64_t count, const mrt_denoiser_train_desc* desc,
                                        mrt_denoiser_trainer** out);
/* The training pairs: count frames of width x height pixels, inputs (10 FP32
 * per pixel) and targets (3 FP32 per pixel), copied to the device (host or
 * device memory per mem).  Frames must be at least crop x crop. */
MRT_API int mrt_denoiser_trainer_set_pairs(mrt_denoiser_trainer* tr, const float* inputs, const float* targets,
                                           uint32_t count, uint32_t width, uint32_t height, int mem);
/* The crops step `step` takes: batch x 5 values (frame, top, left, mirrored
 * left-right, mirrored top-bottom).  Crop i of step s: with mix(z) = the
 * splitmix64 finaliser of z + 0x9E3779B97F4A7C15 and k = s * batch + i,
 * r1 = mix(seed ^ mix(2k)), r2 = mix(seed ^ mix(2k + 1)); frame = r1 mod
 * count, top = (r1 >> 32) mod (height - crop + 1), left = r2 mod (width -
 * crop + 1), mirrored left-right = bit 40 of r2, top-bottom = bit 41. */
MRT_API int mrt_denoiser_trainer_crops(const mrt_denoiser_trainer* tr, uint64_t step, uint32_t* crops);
/* `steps` training steps at learning rate lr; *loss (may be NULL) = the mean
 * of their losses. */
MRT_API int mrt_denoiser_trainer_step(mrt_denoiser_trainer* tr, uint32_t steps, double lr, double* loss);
/* One forward and backward pass over a given batch (batch <= the trainer's,
 * crops of the trainer's side: inputs batch x crop x crop x 10, targets
 * batch x crop x crop x 3, host or device memory per mem): *loss (may be
 * NULL) = its loss, grad (may be NULL, count entries) = the gradient in the
 * weights' layout; with lr > 0 an Adam update follows (a step). */
MRT_API int mrt_denoiser_trainer_gradient(mrt_denoiser_trainer* tr, const float* inputs, const float* targets,
                                          uint32_t batch, double lr, double* loss, float* grad, uint64_t count,
                                          int mem);
/* The FP32 master weights in mrt_denoiser_create's layout. */
MRT_API int mrt_denoiser_trainer_weights(mrt_denoiser_trainer* tr, float* weights, uint64_t count);
MRT_API int mrt_denoiser_trainer_stats(const mrt_denoiser_trainer* tr, mrt_denoiser_train_stats* stats);
MRT_API int mrt_denoiser_trainer_destroy(mrt_denoiser_trainer* tr);

/* ---- host-side construction helpers (no device work) ------------------- */
/* Triangle constructors (Triangle.cpp:12-55): normals/colors may be NULL. */
MRT_API int mrt_triangle_make(const float a[3], const float b[3], const float c[3], int32_t material_id,
                              const float* normals9, const float* colors9, mrt_triangle* out);
/* Camera constructor basis (Camera.cpp:47-54) and Perspective/Orthographic fields. */
MRT_API int mrt_camera_make(int kind, const float position[3], const float look_at[3], const float up[3],
                            float fov_h_or_size_h, float fov_v_or_size_v, mrt_camera* out);
/* Camera::getAABB, the camera's share of the RegularGrid bounds
 * (Shader.cpp:52-62): the position for a perspective camera (Camera.cpp:39-45);
 * for an orthographic one the view rectangle widened by half a unit, with the
 * reference's vertical order (its min takes the top edge, its max the bottom
 * one, Orthographic.cpp:33-45).  out: min xyz, max xyz. */
MRT_API int mrt_camera_aabb(const mrt_camera* camera, float out[6]);
/* Camera tile table (Camera.cpp:16-26): haltonSequence(i, 2) for i < 256,
 * shuffled by a seeded Fisher-Yates (seed 0 keeps the unshuffled order). */
MRT_API int mrt_tile_values(uint32_t seed, float out[256]);
/* Triangle::intersectBox (Triangle.cpp:154-246), used by the grid builder. */
MRT_API int mrt_triangle_intersect_box(const mrt_triangle* tri, const float box_min[3], const float box_max[3]);
/* intersectBox of one primitive of any category: Triangle.cpp:154-246,
 * Sphere.cpp:102-123, Plane.cpp:123-155.  1 overlap, 0 none, < 0 error. */
MRT_API int mrt_primitive_intersect_box(int category, const void* prim, const float box_min[3],
                                        const float box_max[3]);
/* getAABB of n primitives of one category (Triangle.cpp:145-152,
 * Sphere.cpp:91-96, Plane.cpp:80-104): 6 floats (min xyz, max xyz) each. */
MRT_API int mrt_primitive_aabbs(int category, const void* prims, uint32_t n, float* boxes6);
/* BVHNode, app/MobileRT/Accelerators/BVH.hpp:21-25 (32 bytes). */
typedef struct mrt_bvh_node {
    float min[3], max[3];
    uint32_t index_offset;
    uint32_t number_primitives;
} mrt_bvh_node;
/* The reference BVH builder (BVH.hpp:64-155, 274-311) over primitive boxes in
 * scene order; the device layout mrt_scene_build uploads is derived from this
 * node array.  Writes up to `capacity` nodes and the node count. */
MRT_API int mrt_bvh_build(const float* boxes6, uint32_t n, uint32_t leaf_size, mrt_bvh_node* nodes,
                          uint32_t capacity, uint32_t* count);
/* The same builder on the GPU (no counterpart in the reference: its build is
 * sequential): the node array of mrt_bvh_build, bit for bit, built level by
 * level (segmented scans of the box unions, the SAH cost of every split
 * position in parallel, one atomic minimum per node) with subtrees of at
 * most 64 primitives built by one thread each.  boxes6 and nodes in host or
 * device memory per mem; up to `capacity` nodes are written (nodes may be
 * NULL) and *count is the array's length.  MRT_E_UNSUPPORTED when a box has
 * a NaN coordinate or the tree has more than 1024 levels of nodes above 64
 * primitives (mrt_bvh_build covers both). */
MRT_API int mrt_bvh_build_gpu(mrt_ctx* ctx, const float* boxes6, uint32_t n, uint32_t leaf_size,
                              mrt_bvh_node* nodes, uint32_t capacity, uint32_t* count, int mem);
/* RegularGrid cell registration on the GPU (no counterpart in the reference:
 * its grid build is sequential): the cell lists of the grid mrt_scene_build
 * makes (RegularGrid.hpp:50-154: every primitive in the cells of its box
 * range whose box it overlaps by intersectBox), identical to the host build:
 * offsets (grid_size^3 + 1 entries, cell index x + y*gs + z*gs*gs) and refs
 * (ascending primitive indices per cell; up to refs_capacity written, refs
 * may be NULL), *n_refs their total.  prims: n records of the category's
 * struct (mrt_plane, mrt_sphere, mrt_triangle); prims, offsets and refs in
 * host or device memory per mem; bounds6 (min xyz, max xyz) on the host.
 * MRT_E_UNSUPPORTED when a primitive box has a non-finite coordinate. */
MRT_API int mrt_grid_build_gpu(mrt_ctx* ctx, int category, const void* prims, uint32_t n, const float bounds6[6],
                               int32_t grid_size, uint32_t* offsets, uint32_t* refs, uint32_t refs_capacity,
                               uint32_t* n_refs, int mem);
/* The 4-wide tree the persistent trace walks above a set of leaves (box and
 * non-zero leaf word each; scene builds pass the reference BVH's leaves), as
 * built for rays with a finite inverse direction: 32 floats per node (min x,
 * y, z and max x, y, z of its 4 slots as 4-vectors, then 4 (ref, word) pairs
 * as raw bits; a leaf slot holds (leaf index, word), an interior slot (node,
 * 0), an unused slot all-NaN boxes); node 0 holds the top slots; *depth is the
 * number of node levels below node 0.  No counterpart in the reference: a
 * check on the tree's structure. */
MRT_API int mrt_wide_build(const float* boxes6, const uint32_t* words, uint32_t n, float* nodes, uint32_t capacity,
                           uint32_t* count, uint32_t* depth);

/* ---- OBJ / MTL ingestion --------------------------------------------------- */
/* OBJLoader::process + fillScene (app/Components/ObjectLoaders/OBJLoader.cpp:
 * 23-250) over this library's own parser (the reference uses tinyobjloader):
 * X negated, specular halved, transmittance * (1 - dissolve), emission
 * normalised by its largest component when above 1, emissive faces as
 * AreaLights, other materials de-duplicated within Epsilon, polygons split
 * into fans.  mtl_path may be NULL (no materials). */
typedef struct mrt_obj mrt_obj;
MRT_API int mrt_obj_load(const char* obj_path, const char* mtl_path, mrt_obj** out);
MRT_API int mrt_obj_counts(const mrt_obj* obj, uint32_t* n_triangles, uint32_t* n_materials, uint32_t* n_lights,
                           uint32_t* n_faces_skipped);
/* Copies the scene out; every AreaLight gets light_sampler (NULL: RANDOM)
 * with its seed advanced by the light's index. */
MRT_API int mrt_obj_fill(const mrt_obj* obj, mrt_triangle* triangles, mrt_material* materials, mrt_light* lights,
                         const mrt_sampler* light_sampler);
MRT_API int mrt_obj_destroy(mrt_obj* obj);

/* ---- synthetic scenes ---------------------------------------------------- */
/* Procedural Cornell box: 10 wall triangles + 24 box triangles, 2 area lights. */
#define MRT_CORNELL_TRIANGLES 34u
#define MRT_CORNELL_MATERIALS 7u
#define MRT_CORNELL_LIGHTS 2u
MRT_API int mrt_scene_cornell(uint32_t seed, mrt_triangle* tris, mrt_material* mats, mrt_light* lights);
/* Random soup of n triangles in the unit cube (Morton ordered), n_emissive of
 * them emitted as area lights with radiance le; the rest go to tris
 * (n - n_emissive entries) with vertex colours U(0.2,0.8) and material -1. */
MRT_API int mrt_scene_soup(uint32_t seed, uint32_t n, uint32_t n_emissive, float le, float half_extent,
                           mrt_triangle* tris, mrt_light* lights, int light_sampler_kind);
/* Cluster of n_spheres displaced icospheres (icosahedron subdivided subdiv
 * times: 20 * 4^subdiv triangles each) in the unit cube, Morton ordered;
 * n_emissive of the triangles become area lights (outputs as for
 * mrt_scene_soup, total = n_spheres * 20 * 4^subdiv). */
MRT_API int mrt_scene_icospheres(uint32_t seed, uint32_t n_spheres, uint32_t subdiv, float displace,
                                 uint32_t n_emissive, float le, mrt_triangle* tris, mrt_light* lights,
                                 int light_sampler_kind);

#ifdef __cplusplus
}
#endif

#endif /* MOBILERT_GPU_H */
