// Checks of the C++ engine API (include/mobilert/*.hpp), in the style of the
// reference's unit tests.  Modes:
//   api_check host                      host-side behaviour only (no GPU)
//   api_check render <out> <accel>      Cornell Whitted 64x64 through
//                                       Renderer::renderFrame; bitmap to <out>
//   api_check raytrace <out>            the same frame through RayTrace
//   api_check shader <kind> <out> [rt]  the frame with NoShadows / DepthMap /
//                                       DiffuseMaterial (rt: through RayTrace)
//   api_check obj <obj> <mtl>           Components::OBJLoader on the fixture
//   api_check objrt <obj> <mtl> <out>   RayTrace of an OBJ scene
//   api_check rays                      per-ray methods against the batch C ABI
// Prints "ok" and exits 0 on success; prints the failed check and exits 1.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mobilert/Components.hpp"
#include "../../include/mobilert/c_wrapper.h"

using MobileRT::vec3;

static int g_failed = 0;
#define CHECK(cond)                                                               \
    do {                                                                          \
        if (!(cond)) {                                                            \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
            ++g_failed;                                                           \
        }                                                                         \
    } while (0)

static MobileRT::Material mat_of(const mrt_material& m) {
    return MobileRT::Material{vec3{m.kd[0], m.kd[1], m.kd[2]}, vec3{m.ks[0], m.ks[1], m.ks[2]},
                              vec3{m.kt[0], m.kt[1], m.kt[2]}, m.ior, vec3{m.le[0], m.le[1], m.le[2]}};
}

// The Cornell scene of the C ABI generator as a MobileRT::Scene with
// Constant(0.5) light samplers.
static MobileRT::Scene cornell() {
    std::vector<mrt_triangle> t(34);
    std::vector<mrt_material> m(7);
    std::vector<mrt_light> l(2);
    mrt_scene_cornell(1, t.data(), m.data(), l.data());
    MobileRT::Scene s;
    for (const auto& x : t) s.triangles_.emplace_back(x);
    for (const auto& x : m) s.materials_.push_back(mat_of(x));
    for (const auto& x : l)
        s.lights_.push_back(std::make_unique<Components::AreaLight>(
            mat_of(x.radiance), std::make_unique<Components::Constant>(0.5f), MobileRT::Triangle{x.triangle}));
    return s;
}

static std::unique_ptr<MobileRT::Camera> cornell_camera() {
    return std::make_unique<Components::Perspective>(vec3{0.0f, 0.0f, -3.4f}, vec3{0.0f, 0.0f, 1.0f},
                                                     vec3{0.0f, 1.0f, 0.0f}, 45.0f, 45.0f);
}

static int host_checks() {
    // Sampler domain (Sampler.cpp:9-13) and HaltonSeq counter (HaltonSeq.cpp:14-22).
    Components::HaltonSeq h{256, 256, 1};
    CHECK(h.domainSize_ == 256u);
    CHECK(h.getSample(0) == 0.0f);
    CHECK(h.getSample(0) == 0.5f);
    CHECK(h.getSample(0) == 0.25f);
    Components::HaltonSeq hw{1920, 1080, 1};
    CHECK(hw.domainSize_ == 144u);  // width used twice
    h.resetSampling();
    CHECK(h.getSample() == 0.0f);
    Components::Stratified st{256, 256, 1};
    CHECK(st.getSample(0) == 0.0f);
    CHECK(st.getSample(0) == 1.0f / 256.0f);
    // Past the domain (256 draws at 16x16) every value is 1.0.
    Components::Stratified tiny{16, 16, 1};
    CHECK(tiny.domainSize_ == 256u);
    for (int i = 0; i < 256; ++i) tiny.getSample(0);
    CHECK(tiny.getSample(0) == 1.0f);
    CHECK(tiny.getSample(1) == 0.0f);
    Components::Constant c{0.25f};
    CHECK(c.getSample(7) == 0.25f);
    CHECK(c.descriptor(true).kind == MRT_SAMPLER_CONSTANT);
    CHECK(h.descriptor(true).kind == MRT_SAMPLER_HALTON_COMPAT && h.descriptor(true).domain_size == 256u);
    CHECK(h.descriptor(false).kind == MRT_SAMPLER_HALTON);
    Components::MersenneTwister mt1{5}, mt2{5};
    for (int i = 0; i < 8; ++i) {
        const float a = mt1.getSample(), b = mt2.getSample();
        CHECK(a == b && a >= 0.0f && a < 1.0f);
    }
    // Camera basis (Camera.cpp:47-54) and the tile scheduler (Camera.cpp:28-37).
    auto cam = cornell_camera();
    CHECK(cam->direction_ == (vec3{0.0f, 0.0f, 1.0f}));
    CHECK(cam->right_ == (vec3{1.0f, 0.0f, 0.0f}));
    CHECK(cam->up_ == (vec3{0.0f, 1.0f, 0.0f}));
    std::vector<int> seen(256, 0);
    for (int i = 0; i < 256; ++i) {
        const float b = cam->getBlock(0);
        CHECK(b < 1.0f);
        const int k = static_cast<int>(std::lround(b * 256.0f));
        if (k >= 0 && k < 256) ++seen[k];
    }
    CHECK(cam->getBlock(0) == 1.0f);
    for (int v : seen) CHECK(v == 1);
    CHECK(cam->getBlock(1) < 1.0f);
    cam->resetSampling();
    CHECK(cam->getBlock(0) == 0.0f);  // seed 0 keeps the Halton order
    // Triangle construction (Triangle.cpp:12-55) and its box.
    MobileRT::Triangle t{vec3{0, 0, 0}, vec3{1, 0, 0}, vec3{0, 1, 0}, 3};
    CHECK(t.AB_ == (vec3{1, 0, 0}) && t.AC_ == (vec3{0, 1, 0}) && t.materialId_ == 3);
    const MobileRT::AABB b = t.getAABB();
    CHECK(b.pointMin_ == (vec3{0, 0, 0}) && b.pointMax_ == (vec3{1, 1, 0}));
    CHECK(t.intersectBox(MobileRT::AABB{vec3{-1, -1, -1}, vec3{2, 2, 2}}));
    const MobileRT::AABB u = MobileRT::surroundingBox(b, MobileRT::AABB{vec3{-1, 0, 0}, vec3{0, 0, 5}});
    CHECK(u.pointMin_ == (vec3{-1, 0, 0}) && u.pointMax_ == (vec3{1, 1, 5}));
    CHECK(MobileRT::AABB(vec3{0, 0, 0}, vec3{1, 2, 3}).getSurfaceArea() == 22.0f);
    // Material equality within Epsilon (Material.cpp:21-29).
    const MobileRT::Material m1{vec3{0.5f}}, m2{vec3{0.5f + 1e-7f}}, m3{vec3{0.6f}};
    CHECK(m1 == m2);
    CHECK(!(m1 == m3));
    // Sphere box and plane distance.
    MobileRT::Sphere sp{vec3{1, 2, 3}, 0.5f, 0};
    CHECK(sp.getAABB().pointMin_ == (vec3{0.5f, 1.5f, 2.5f}));
    CHECK(sp.getZ() == 3.0f);
    MobileRT::Plane pl{vec3{0, -1, 0}, vec3{0, 1, 0}, 0};
    CHECK(pl.distance(vec3{0, 1, 0}) == 2.0f);
    return g_failed;
}

static int render(const char* out, int accel, bool viaRayTrace) {
    const uint32_t w = 64, h = 64;
    std::vector<uint32_t> bm(w * h, 0u);
    if (viaRayTrace) {
        RayTrace(bm.data(), static_cast<int32_t>(w), static_cast<int32_t>(h), 4, MRT_SHADER_WHITTED, 0, 1, 1, 1,
                 accel, false, false, nullptr, nullptr);
    } else {
        auto shader = std::make_unique<Components::Whitted>(cornell(), 1u, static_cast<MobileRT::Shader::Accelerator>(accel));
        MobileRT::Renderer r{std::move(shader), cornell_camera(), std::make_unique<Components::Constant>(0.5f), w, h, 1};
        // Accelerators own the primitives after construction (Shader.cpp:34-36).
        CHECK(r.shader_->scene_.triangles_.empty());
        r.renderFrame(bm.data(), 8, w * 4u);
        CHECK(r.getSample() == 1u);
        CHECK(r.lastStats().rays_camera == w * h);
    }
    FILE* f = std::fopen(out, "wb");
    if (!f) return 1;
    std::fwrite(bm.data(), 4, bm.size(), f);
    std::fclose(f);
    return g_failed;
}

static int render_shader(int kind, const char* out, bool viaRayTrace) {
    const uint32_t w = 64, h = 64;
    std::vector<uint32_t> bm(w * h, 0u);
    if (viaRayTrace) {
        RayTrace(bm.data(), static_cast<int32_t>(w), static_cast<int32_t>(h), 4, kind, 0, 1, 1, 1, MRT_BVH, false,
                 false, nullptr, nullptr);
    } else {
        std::unique_ptr<MobileRT::Shader> shader;
        if (kind == MRT_SHADER_DEPTHMAP)
            shader = std::make_unique<Components::DepthMap>(cornell(), vec3{1, 1, 1}, MobileRT::Shader::BVH);
        else if (kind == MRT_SHADER_DIFFUSE)
            shader = std::make_unique<Components::DiffuseMaterial>(cornell(), MobileRT::Shader::BVH);
        else
            shader = std::make_unique<Components::NoShadows>(cornell(), 1u, MobileRT::Shader::BVH);
        MobileRT::Renderer r{std::move(shader), cornell_camera(), std::make_unique<Components::Constant>(0.5f), w, h, 1};
        r.renderFrame(bm.data(), 8, w * 4u);
        CHECK(r.lastStats().rays_camera == w * h);
        CHECK(r.lastStats().rays_shadow == 0u);
    }
    FILE* f = std::fopen(out, "wb");
    if (!f) return 1;
    std::fwrite(bm.data(), 4, bm.size(), f);
    std::fclose(f);
    return g_failed;
}

// OBJLoader.cpp:23-250 through the C++ API: counts, and materials
// deduplicated against one the scene already holds.
static int obj(const char* objPath, const char* mtlPath) {
    Components::OBJLoader loader{objPath, mtlPath};
    CHECK(loader.process() == 9);  // 7 triangles + 2 emissive
    CHECK(loader.isProcessed());
    MobileRT::Scene scene;
    // The fixture's "mirror" after conversion: Ks halved, ior 1.5.
    scene.materials_.push_back(MobileRT::Material{vec3{0.0f}, vec3{0.45f, 0.4f, 0.35f}, vec3{0.0f}, 1.5f});
    int made = 0;
    CHECK(loader.fillScene(&scene, [&made] {
        ++made;
        return std::make_unique<Components::Constant>(0.5f);
    }));
    CHECK(made == 2 && scene.lights_.size() == 2u);
    CHECK(scene.triangles_.size() == 7u && scene.materials_.size() == 4u);
    const int expect[7] = {1, 1, 0, 0, 2, -1, 3};
    for (int k = 0; k < 7; ++k) CHECK(scene.triangles_[k].materialId_ == expect[k]);
    CHECK(scene.lights_[0]->radiance_.Le_ == (vec3{1.0f, 0.5f, 0.25f}));
    Components::OBJLoader missing{"no_such.obj", ""};
    CHECK(missing.process() == -1 && !missing.isProcessed());
    return g_failed;
}

static int obj_raytrace(const char* objPath, const char* mtlPath, const char* out) {
    const uint32_t w = 64, h = 64;
    std::vector<uint32_t> bm(w * h, 0u);
    RayTrace(bm.data(), static_cast<int32_t>(w), static_cast<int32_t>(h), 4, MRT_SHADER_WHITTED, 5, 1, 1, 1, MRT_BVH,
             false, false, objPath, mtlPath);
    int lit = 0;
    for (uint32_t v : bm) lit += (v & 0xFFFFFFu) != 0u;
    CHECK(lit > 0);
    FILE* f = std::fopen(out, "wb");
    if (!f) return 1;
    std::fwrite(bm.data(), 4, bm.size(), f);
    std::fclose(f);
    return g_failed;
}

// Per-ray methods against the batch C ABI on the same inputs.
static int rays() {
    auto shader = std::make_unique<Components::Whitted>(cornell(), 1u, MobileRT::Shader::BVH);
    auto cam = cornell_camera();
    shader->initializeAccelerators(cam.get());
    const auto& tris = shader->bvhTriangles_.primitives_;
    CHECK(tris.size() == 34u);
    int hits = 0;
    for (int k = 0; k < 64; ++k) {
        const float u = (k % 8) / 8.0f + 0.03f, v = (k / 8) / 8.0f + 0.05f;
        const MobileRT::Ray ray = cam->generateRay(u, v, 0.0f, 0.0f);
        // Accelerator closest hit against the brute-force Triangle tests.
        MobileRT::Intersection a{MobileRT::RayLengthMax};
        const bool ha = shader->bvhTriangles_.trace(&a, ray);
        MobileRT::Intersection b{MobileRT::RayLengthMax};
        bool hb = false;
        for (const auto& t : tris) hb = t.intersect(&b, ray) || hb;
        CHECK(ha == hb);
        if (ha && hb) {
            ++hits;
            CHECK(a.length_ == b.length_ && a.primitive_ == b.primitive_);
            CHECK(a.normal_ == b.normal_ && a.point_ == b.point_);
        }
        // A shadow ray from the hit towards the light, any-hit vs closest.
        if (ha) {
            const MobileRT::Ray up{vec3{0.0f, 1.0f, 0.0f}, a.point_, 2, a.primitive_};
            MobileRT::Intersection c{MobileRT::RayLengthMax};
            const bool closest = shader->bvhTriangles_.trace(&c, up);
            CHECK(shader->bvhTriangles_.shadowTrace(up, 1.0e30f) == closest);
            CHECK(shader->shadowTrace(up, 1.0e30f) == closest);
            CHECK(!shader->shadowTrace(up, 1.0e-7f));
        }
        // rayTrace radiance is finite and non-negative.
        vec3 rgb{};
        shader->rayTrace(&rgb, ray);
        CHECK(std::isfinite(rgb.x) && rgb.x >= 0.0f && rgb.y >= 0.0f && rgb.z >= 0.0f);
    }
    CHECK(hits > 32);
    // Box test on the device.
    const MobileRT::Ray r0{vec3{0.0f, 0.0f, 1.0f}, vec3{0.0f, 0.0f, -5.0f}, 1};
    CHECK(MobileRT::intersect(MobileRT::AABB{vec3{-1, -1, -1}, vec3{1, 1, 1}}, r0));
    CHECK(!MobileRT::intersect(MobileRT::AABB{vec3{2, 2, -1}, vec3{3, 3, 1}}, r0));
    // The light is hit by traceLights, not by the triangle accelerator.
    const MobileRT::Ray toLight{vec3{0.0f, 1.0f, 0.0f}, vec3{0.0f, 0.0f, 0.0f}, 1};
    MobileRT::Intersection li{MobileRT::RayLengthMax};
    CHECK(shader->scene_.traceLights(&li, toLight));
    CHECK(li.material_ != nullptr && li.material_->Le_.x > 0.0f);
    return g_failed;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string mode = argv[1];
    int rc = 2;
    if (mode == "host") rc = host_checks();
    else if (mode == "render" && argc >= 4) rc = render(argv[2], std::atoi(argv[3]), false);
    else if (mode == "raytrace" && argc >= 3) rc = render(argv[2], MRT_BVH, true);
    else if (mode == "rays") rc = rays();
    else if (mode == "shader" && argc >= 4) rc = render_shader(std::atoi(argv[2]), argv[3], argc >= 5);
    else if (mode == "obj" && argc >= 4) rc = obj(argv[2], argv[3]);
    else if (mode == "objrt" && argc >= 5) rc = obj_raytrace(argv[2], argv[3], argv[4]);
    if (rc == 0) std::printf("ok\n");
    return rc == 0 ? 0 : 1;
}
