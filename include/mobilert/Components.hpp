/* MobileRT's plugins (namespace Components) over the MI355X library: the
 * reference's shaders, samplers, cameras and lights (app/Components/...),
 * same names and constructors.  Each plugin also names the device-side
 * equivalent that renders use (descriptor()); see MobileRT.hpp for how the
 * engine classes map onto the C ABI. */
#pragma once

#include <functional>
#include <string>

#include "MobileRT.hpp"

namespace Components {

/* ---- shaders (app/Components/Shaders) ---------------------------------- */

/* Whitted.hpp: direct light from every light sample, reflection and
 * transmission recursion, 0.1 * Kd ambient (Whitted.cpp:17-107). */
class Whitted final : public ::MobileRT::Shader {
private:
    bool shade(::MobileRT::vec3* rgb, const ::MobileRT::Intersection& intersection,
               const ::MobileRT::Ray& ray) noexcept final;
    int kind() const noexcept final { return MRT_SHADER_WHITTED; }

public:
    explicit Whitted(::MobileRT::Scene&& scene, std::uint32_t samplesLight, Accelerator accelerator) noexcept;
};

/* NoShadows.hpp: the direct light of every light sample without shadow
 * rays, plus 0.1 * Kd ambient (NoShadows.cpp:19-64). */
class NoShadows final : public ::MobileRT::Shader {
private:
    bool shade(::MobileRT::vec3* rgb, const ::MobileRT::Intersection& intersection,
               const ::MobileRT::Ray& ray) noexcept final;
    int kind() const noexcept final { return MRT_SHADER_NOSHADOWS; }

public:
    explicit NoShadows(::MobileRT::Scene&& scene, std::uint32_t samplesLight, Accelerator accelerator) noexcept;
};

/* DepthMap.hpp: grey level of the hit distance against 2 |maxPoint - origin|
 * (DepthMap.cpp:19-26). */
class DepthMap final : public ::MobileRT::Shader {
private:
    ::MobileRT::vec3 maxPoint_{};
    bool shade(::MobileRT::vec3* rgb, const ::MobileRT::Intersection& intersection,
               const ::MobileRT::Ray& ray) noexcept final;
    int kind() const noexcept final { return MRT_SHADER_DEPTHMAP; }
    void describeShader(mrt_render_desc* desc) const noexcept final;

public:
    explicit DepthMap(::MobileRT::Scene&& scene, const ::MobileRT::vec3& maxPoint, Accelerator accelerator) noexcept;
};

/* DiffuseMaterial.hpp: the first non-zero of Kd, Ks, Kt, Le, or the
 * primitive colour (DiffuseMaterial.cpp:16-38). */
class DiffuseMaterial final : public ::MobileRT::Shader {
private:
    bool shade(::MobileRT::vec3* rgb, const ::MobileRT::Intersection& intersection,
               const ::MobileRT::Ray& ray) noexcept final;
    int kind() const noexcept final { return MRT_SHADER_DIFFUSE; }

public:
    explicit DiffuseMaterial(::MobileRT::Scene&& scene, Accelerator accelerator) noexcept;
};

/* PathTracer.hpp: one light sample estimate per vertex, cosine-weighted
 * continuation with Russian roulette beyond RayDepthMin
 * (PathTracer.cpp:25-164). */
class PathTracer final : public ::MobileRT::Shader {
private:
    std::unique_ptr<::MobileRT::Sampler> samplerRussianRoulette_{};
    bool shade(::MobileRT::vec3* rgb, const ::MobileRT::Intersection& intersection,
               const ::MobileRT::Ray& ray) noexcept final;
    int kind() const noexcept final { return MRT_SHADER_PATHTRACER; }
    mrt_sampler samplerRussianRoulette() const noexcept final;

public:
    explicit PathTracer(::MobileRT::Scene&& scene, std::unique_ptr<::MobileRT::Sampler> samplerRussianRoulette,
                        std::uint32_t samplesLight, Accelerator accelerator) noexcept;
    void resetSampling() noexcept final;
};

/* ---- samplers (app/Components/Samplers) -------------------------------- */

/* Constant.hpp: always value_. */
class Constant final : public ::MobileRT::Sampler {
public:
    using ::MobileRT::Sampler::getSample;

public:
    const float value_{};

public:
    explicit Constant(float value) noexcept;
    float getSample(std::uint32_t sample) noexcept final;
    mrt_sampler descriptor(bool pixel) const noexcept final;
};

/* HaltonSeq.hpp: base-2 radical inverse of the shared counter
 * (HaltonSeq.cpp:14-22).  As a pixel sampler renders replay that counter in
 * single-thread claim order (MRT_SAMPLER_HALTON_COMPAT); elsewhere, and
 * with perPixel(true), the per-pixel scrambled Halton stream. */
class HaltonSeq final : public ::MobileRT::Sampler {
public:
    using ::MobileRT::Sampler::getSample;

private:
    bool perPixel_{false};
    std::uint32_t seed_{0};

public:
    explicit HaltonSeq(std::uint32_t width, std::uint32_t height, std::uint32_t samples) noexcept;
    float getSample(std::uint32_t sample) noexcept final;
    mrt_sampler descriptor(bool pixel) const noexcept final;
    /* Build-specific: per-pixel scrambled Halton draws with this seed. */
    void perPixel(bool on, std::uint32_t seed = 0) noexcept { perPixel_ = on; seed_ = seed; }
};

/* Stratified.hpp: (c - s * D) / D of the shared counter
 * (Stratified.cpp:14-23); pixel renders replay it like HaltonSeq. */
class Stratified final : public ::MobileRT::Sampler {
public:
    using ::MobileRT::Sampler::getSample;

private:
    bool perPixel_{false};
    std::uint32_t seed_{0};

public:
    explicit Stratified(std::uint32_t width, std::uint32_t height, std::uint32_t samples) noexcept;
    float getSample(std::uint32_t sample) noexcept final;
    mrt_sampler descriptor(bool pixel) const noexcept final;
    void perPixel(bool on, std::uint32_t seed = 0) noexcept { perPixel_ = on; seed_ = seed; }
};

/* MersenneTwister.hpp: uniform [0, 1) draws.  The reference seeds a
 * thread-local mt19937 from /dev/urandom; here the stream is the seeded
 * counter hash of the deterministic contract (a fresh seed per instance
 * unless one is given). */
class MersenneTwister final : public ::MobileRT::Sampler {
public:
    using ::MobileRT::Sampler::getSample;

private:
    std::uint32_t seed_{};

public:
    explicit MersenneTwister() noexcept;
    explicit MersenneTwister(std::uint32_t seed) noexcept;
    float getSample(std::uint32_t sample) noexcept final;
    mrt_sampler descriptor(bool pixel) const noexcept final;
};

/* ---- cameras (app/Components/Cameras) ---------------------------------- */

/* Perspective.hpp (Perspective.cpp:24-44); fovs in degrees. */
class Perspective final : public ::MobileRT::Camera {
private:
    float hFov_{};
    float vFov_{};

public:
    explicit Perspective(const ::MobileRT::vec3& position, const ::MobileRT::vec3& lookAt,
                         const ::MobileRT::vec3& up, float hFov, float vFov) noexcept;
    ::MobileRT::Ray generateRay(float u, float v, float deviationU, float deviationV) const noexcept final;
    float getHFov() const noexcept;
    float getVFov() const noexcept;
    mrt_camera descriptor() const noexcept final;
};

/* Orthographic.hpp (Orthographic.cpp:23-45). */
class Orthographic final : public ::MobileRT::Camera {
private:
    float sizeH_{};
    float sizeV_{};

public:
    explicit Orthographic(const ::MobileRT::vec3& position, const ::MobileRT::vec3& lookAt,
                          const ::MobileRT::vec3& up, float sizeH, float sizeV) noexcept;
    ::MobileRT::Ray generateRay(float u, float v, float deviationU, float deviationV) const noexcept final;
    ::MobileRT::AABB getAABB() const noexcept final;
    float getSizeH() const noexcept;
    float getSizeV() const noexcept;
    mrt_camera descriptor() const noexcept final;
};

/* ---- lights (app/Components/Lights) ------------------------------------ */

/* AreaLight.hpp: an emissive triangle; getPosition draws (R, S) from its
 * sampler and folds R + S >= 1 (AreaLight.cpp:22-31). */
class AreaLight final : public ::MobileRT::Light {
private:
    ::MobileRT::Triangle triangle_;
    std::unique_ptr<::MobileRT::Sampler> samplerPointLight_{};

public:
    explicit AreaLight(const ::MobileRT::Material& radiance, std::unique_ptr<::MobileRT::Sampler> samplerPointLight,
                       const ::MobileRT::vec3& pointA, const ::MobileRT::vec3& pointB,
                       const ::MobileRT::vec3& pointC) noexcept;
    /* Build-specific: a light over an existing triangle (its AB / AC kept
     * bit for bit, as scene generators emit them). */
    explicit AreaLight(const ::MobileRT::Material& radiance, std::unique_ptr<::MobileRT::Sampler> samplerPointLight,
                       const ::MobileRT::Triangle& triangle) noexcept;
    ::MobileRT::vec3 getPosition() noexcept final;
    void resetSampling() noexcept final;
    bool intersect(::MobileRT::Intersection* intersection, const ::MobileRT::Ray& ray) const noexcept final;
    mrt_light descriptor() const noexcept final;
    const ::MobileRT::Triangle& triangle() const noexcept { return triangle_; }
};

/* PointLight.hpp: a fixed position that rays never hit. */
class PointLight final : public ::MobileRT::Light {
private:
    ::MobileRT::vec3 position_{};

public:
    explicit PointLight(const ::MobileRT::Material& radiance, const ::MobileRT::vec3& position) noexcept;
    ::MobileRT::vec3 getPosition() noexcept final;
    void resetSampling() noexcept final;
    bool intersect(::MobileRT::Intersection* intersection, const ::MobileRT::Ray& ray) const noexcept final;
    mrt_light descriptor() const noexcept final;
};

/* ---- object loaders (app/Components/ObjectLoaders) --------------------- */

/* OBJLoader.hpp: OBJ / MTL files into a Scene (OBJLoader.cpp:23-250), over
 * this library's parser (mrt_obj_load; the reference uses tinyobjloader). */
class OBJLoader final {
private:
    std::string objFilePath_{};
    std::string mtlFilePath_{};
    mrt_obj* obj_{nullptr};
    bool isProcessed_{false};
    std::int32_t numberTriangles_{0};

public:
    explicit OBJLoader(std::string obj, std::string materials) noexcept;
    OBJLoader(const OBJLoader&) = delete;
    OBJLoader& operator=(const OBJLoader&) = delete;
    ~OBJLoader() noexcept;
    /* Parses the files: the number of triangles (emissive ones included),
     * -1 on failure. */
    std::int32_t process() noexcept;
    bool isProcessed() const noexcept { return isProcessed_; }
    /* Appends the triangles, materials (each deduplicated against the
     * scene's, Material.cpp:21-29) and AreaLights (each with a sampler from
     * lambda) to the scene. */
    bool fillScene(::MobileRT::Scene* scene, std::function<std::unique_ptr<::MobileRT::Sampler>()> lambda) noexcept;
};

}  // namespace Components
