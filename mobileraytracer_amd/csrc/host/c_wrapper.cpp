// RayTrace, the reference's outer C entry point (c_wrapper.h:10-15 and
// c_wrapper.cpp:25-548), over the C++ API: build a scene, camera, pixel
// sampler and shader, then render `repeats` frames into the caller's bitmap.
#include <chrono>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/mobilert/Components.hpp"
#include "../../../include/mobilert/c_wrapper.h"

namespace {

using MobileRT::vec3;

std::unique_ptr<MobileRT::Sampler> sampler_from(const mrt_sampler& s) {
    if (s.kind == MRT_SAMPLER_RANDOM) return std::make_unique<Components::MersenneTwister>(s.seed);
    return std::make_unique<Components::Constant>(s.value);
}

MobileRT::Material material_from(const mrt_material& m) {
    return MobileRT::Material{vec3{m.kd[0], m.kd[1], m.kd[2]}, vec3{m.ks[0], m.ks[1], m.ks[2]},
                              vec3{m.kt[0], m.kt[1], m.kt[2]}, m.ior, vec3{m.le[0], m.le[1], m.le[2]}};
}

// The scene generators of the C ABI as a MobileRT::Scene.
bool make_scene(int32_t id, MobileRT::Scene* scene, std::unique_ptr<MobileRT::Camera>* camera, float hfov,
                float vfov) {
    std::vector<mrt_triangle> tris;
    std::vector<mrt_material> mats;
    std::vector<mrt_light> lights;
    if (id == 0) {
        tris.resize(34);
        mats.resize(7);
        lights.resize(2);
        if (mrt_scene_cornell(1, tris.data(), mats.data(), lights.data()) != MRT_OK) return false;
        *camera = std::make_unique<Components::Perspective>(vec3{0.0f, 0.0f, -3.4f}, vec3{0.0f, 0.0f, 1.0f},
                                                            vec3{0.0f, 1.0f, 0.0f}, hfov, vfov);
    } else if (id == 1) {
        const uint32_t n = 1000000, ne = 10000;
        tris.resize(n - ne);
        lights.resize(ne);
        if (mrt_scene_soup(2, n, ne, 5.0f, 0.005f, tris.data(), lights.data(), MRT_SAMPLER_RANDOM) != MRT_OK)
            return false;
        *camera = std::make_unique<Components::Perspective>(vec3{0.5f, 0.5f, -1.0f}, vec3{0.5f, 0.5f, 0.5f},
                                                            vec3{0.0f, 1.0f, 0.0f}, hfov, vfov);
    } else {
        return false;
    }
    scene->triangles_.reserve(tris.size());
    for (const mrt_triangle& t : tris) scene->triangles_.emplace_back(t);
    for (const mrt_material& m : mats) scene->materials_.push_back(material_from(m));
    for (const mrt_light& l : lights)
        scene->lights_.push_back(std::make_unique<Components::AreaLight>(material_from(l.radiance),
                                                                         sampler_from(l.sampler),
                                                                         MobileRT::Triangle{l.triangle}));
    return true;
}

void work(uint32_t* bitmap, int32_t width, int32_t height, int32_t threads, int32_t nShader, int32_t nScene,
          int32_t samplesPixel, int32_t samplesLight, int32_t repeats, int32_t accelerator, bool printStdOut,
          std::string pathObj, std::string pathMtl) {
    if (accelerator < 0 || accelerator > 2 || width <= 0 || height <= 0 || samplesPixel <= 0 || samplesLight <= 0) {
        std::fprintf(stderr, "RayTrace: invalid arguments\n");
        return;
    }
    // c_wrapper.cpp:56-58: 45 degrees on the shorter side.
    const float ratio = static_cast<float>(width) / static_cast<float>(height);
    const float hfov = 45.0f * std::max(ratio, 1.0f);
    const float vfov = 45.0f * std::max(1.0f / ratio, 1.0f);
    const auto t0 = std::chrono::steady_clock::now();
    MobileRT::Scene scene;
    std::unique_ptr<MobileRT::Camera> camera;
    if (nScene != 0 && nScene != 1 && !pathObj.empty()) {
        // c_wrapper.cpp:115-135: the OBJ scene (its emissive faces as
        // MersenneTwister-sampled AreaLights) with the default camera
        // (c_wrapper.cpp:62-67), or the CornellBox files' pose.  The other
        // named-file presets (extra lights and poses for the reference's demo
        // models) are not reproduced.
        Components::OBJLoader loader{pathObj, pathMtl};
        loader.process();
        if (!loader.isProcessed() ||
            !loader.fillScene(&scene, [] { return std::make_unique<Components::MersenneTwister>(); })) {
            std::fprintf(stderr, "RayTrace: cannot load %s\n", pathObj.c_str());
            return;
        }
        if (pathObj.find("CornellBox") != std::string::npos)
            camera = std::make_unique<Components::Perspective>(vec3{0.0f, 0.7f, 3.0f}, vec3{0.0f, 0.7f, -1.0f},
                                                               vec3{0.0f, 1.0f, 0.0f}, hfov, vfov);
        else
            camera = std::make_unique<Components::Perspective>(vec3{0.0f, 0.0f, -3.4f}, vec3{0.0f, 0.0f, 1.0f},
                                                               vec3{0.0f, 1.0f, 0.0f}, hfov, vfov);
    } else if (!make_scene(nScene, &scene, &camera, hfov, vfov)) {
        std::fprintf(stderr, "RayTrace: scene %d is not available (0 Cornell box, 1 1M-triangle soup, other ids "
                             "with an OBJ path)\n", nScene);
        return;
    }
    const size_t nTriangles = scene.triangles_.size(), nLights = scene.lights_.size();
    // c_wrapper.cpp:431-438.
    std::unique_ptr<MobileRT::Sampler> samplerPixel;
    if (samplesPixel > 1) samplerPixel = std::make_unique<Components::MersenneTwister>();
    else samplerPixel = std::make_unique<Components::Constant>(0.5f);
    const auto acc = static_cast<MobileRT::Shader::Accelerator>(accelerator);
    std::unique_ptr<MobileRT::Shader> shader;
    // c_wrapper.cpp:444-486: 1 Whitted, 2 PathTracer, 3 DepthMap, 4
    // DiffuseMaterial, anything else NoShadows; maxDist (1, 1, 1) for the
    // unit-sized scenes (c_wrapper.cpp:78).
    if (nShader == MRT_SHADER_WHITTED)
        shader = std::make_unique<Components::Whitted>(std::move(scene), static_cast<uint32_t>(samplesLight), acc);
    else if (nShader == MRT_SHADER_PATHTRACER)
        shader = std::make_unique<Components::PathTracer>(std::move(scene),
                                                          std::make_unique<Components::MersenneTwister>(),
                                                          static_cast<uint32_t>(samplesLight), acc);
    else if (nShader == MRT_SHADER_DEPTHMAP)
        shader = std::make_unique<Components::DepthMap>(std::move(scene), MobileRT::vec3{1.0f, 1.0f, 1.0f}, acc);
    else if (nShader == MRT_SHADER_DIFFUSE)
        shader = std::make_unique<Components::DiffuseMaterial>(std::move(scene), acc);
    else
        shader = std::make_unique<Components::NoShadows>(std::move(scene), static_cast<uint32_t>(samplesLight), acc);
    MobileRT::Renderer renderer{std::move(shader), std::move(camera), std::move(samplerPixel),
                                static_cast<uint32_t>(width), static_cast<uint32_t>(height),
                                static_cast<uint32_t>(samplesPixel)};
    const auto t1 = std::chrono::steady_clock::now();
    do {
        renderer.renderFrame(bitmap, threads, static_cast<uint32_t>(width) * 4u);
    } while (--repeats > 0);
    const auto t2 = std::chrono::steady_clock::now();
    if (printStdOut) {
        const mrt_render_stats& st = renderer.lastStats();
        std::printf("TRIANGLES = %zu\nLIGHTS = %zu\nnShader = %d\nnScene = %d\nsamplesPixel = %d\n"
                    "samplesLight = %d\nwidth_ = %d\nheight_ = %d\n",
                    nTriangles, nLights, nShader, nScene, samplesPixel, samplesLight, width, height);
        std::printf("Creating Time in secs = %f\nRendering Time in secs = %f\n",
                    std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count());
        std::printf("rays: %llu camera, %llu extension, %llu shadow, %llu specular\n",
                    static_cast<unsigned long long>(st.rays_camera), static_cast<unsigned long long>(st.rays_extension),
                    static_cast<unsigned long long>(st.rays_shadow), static_cast<unsigned long long>(st.rays_specular));
        std::fflush(stdout);
    }
}

}  // namespace

extern "C" void RayTrace(uint32_t* bitmap, int32_t width, int32_t height, int32_t threads, int32_t shader,
                         int32_t scene, int32_t samplesPixel, int32_t samplesLight, int32_t repeats,
                         int32_t accelerator, bool printStdOut, bool async, const char* pathObj,
                         const char* pathMtl) {
    const std::string obj = pathObj ? pathObj : "", mtl = pathMtl ? pathMtl : "";
    if (async) {
        std::thread{work, bitmap, width, height, threads, shader, scene, samplesPixel, samplesLight, repeats,
                    accelerator, printStdOut, obj, mtl}
            .detach();
    } else {
        work(bitmap, width, height, threads, shader, scene, samplesPixel, samplesLight, repeats, accelerator,
             printStdOut, obj, mtl);
    }
}
