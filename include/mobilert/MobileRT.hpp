/* MobileRT's C++ engine API (namespace MobileRT) over the MI355X library.
 *
 * The classes keep the reference's names, members and signatures
 * (the app/MobileRT headers), so code written against the reference engine
 * compiles against this header with its vector type swapped for
 * MobileRT::vec3 (three floats, the layout of glm::vec3; it converts from any
 * type with x, y, z members).  Every computation runs on the GPU through the
 * C ABI of mobilert_gpu.h:
 *   Renderer::renderFrame        -> mrt_render (the whole frame in HBM)
 *   Shader::rayTrace             -> mrt_shade_rays (a one-ray batch)
 *   Shader::shadowTrace          -> mrt_trace_any
 *   Naive/RegularGrid/BVH::trace -> mrt_trace_closest on the category's own
 *                                   device scene and accelerator
 *   Triangle/Sphere/Plane::intersect, AreaLight::intersect
 *                                -> mrt_intersect_pairs (one-ray batches)
 *   Camera::generateRay          -> mrt_camera_rays
 * Per-ray methods exist for API compatibility and tests; throughput work
 * goes through renderFrame or the batch entry points of the C ABI.
 *
 * Differences a caller can see:
 *  - Failures are reported, not fatal: methods stay noexcept, log to stderr
 *    and leave outputs untouched; MobileRT::lastError() returns the message.
 *  - renderFrame's numThreads is accepted and ignored (the frame runs on the
 *    GPU); width must be a multiple of 16 as in the reference's Linux front
 *    end.
 *  - Random draws (light choice, hemisphere, MersenneTwister, Russian
 *    roulette) follow the seeded deterministic contract instead of
 *    thread-local /dev/urandom generators, so frames are reproducible. */
#pragma once

#include <atomic>
#include <cmath>
#include <cstdint>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "../mobilert_gpu.h"

namespace MobileRT {

/* Reference constants, app/MobileRT/Utils.hpp:33-37. */
constexpr float Epsilon{1.0e-06f};
constexpr float RayLengthMax{1.0e+30f};
constexpr std::int32_t RayDepthMin{4};
constexpr std::int32_t RayDepthMax{8};

/* Last failure message of any call of this API on this thread. */
const std::string& lastError() noexcept;

/* Three floats with glm::vec3's layout and element-wise arithmetic. */
struct vec3 {
    float x{}, y{}, z{};
    constexpr vec3() noexcept = default;
    constexpr explicit vec3(float s) noexcept : x{s}, y{s}, z{s} {}
    constexpr vec3(float a, float b, float c) noexcept : x{a}, y{b}, z{c} {}
    template <class V, class = decltype(std::declval<const V&>().x)>
    constexpr vec3(const V& v) noexcept : x{static_cast<float>(v.x)}, y{static_cast<float>(v.y)},
                                          z{static_cast<float>(v.z)} {}
    constexpr float operator[](int i) const noexcept { return i == 0 ? x : (i == 1 ? y : z); }
    friend constexpr vec3 operator+(vec3 a, vec3 b) noexcept { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
    friend constexpr vec3 operator-(vec3 a, vec3 b) noexcept { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
    friend constexpr vec3 operator*(vec3 a, vec3 b) noexcept { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
    friend constexpr vec3 operator*(vec3 a, float s) noexcept { return {a.x * s, a.y * s, a.z * s}; }
    friend constexpr vec3 operator/(vec3 a, float s) noexcept { return {a.x / s, a.y / s, a.z / s}; }
    friend constexpr bool operator==(vec3 a, vec3 b) noexcept { return a.x == b.x && a.y == b.y && a.z == b.z; }
};

/* app/MobileRT/Ray.hpp:14-17.  primitive_ is the object the ray leaves (a
 * Triangle, Sphere or Plane of the shader's scene, or nullptr). */
class Ray final {
public:
    const vec3 origin_{};
    const vec3 direction_{};
    const std::int32_t depth_{};
    const void* const primitive_{};

public:
    explicit Ray(const vec3& dir, const vec3& origin, std::int32_t depth, const void* primitive = nullptr) noexcept;
};

/* app/MobileRT/Material.hpp. */
class Material final {
public:
    vec3 Le_{};
    vec3 Kd_{};
    vec3 Ks_{};
    vec3 Kt_{};
    float refractiveIndice_{};

public:
    explicit Material(const vec3& Kd, const vec3& Ks = vec3{}, const vec3& Kt = vec3{},
                      float refractiveIndice = 1.0f, const vec3& Le = vec3{}) noexcept;
    /* Material.cpp:21-29: every field within Epsilon. */
    bool operator==(const Material& material) const noexcept;
    mrt_material descriptor() const noexcept;
};

/* app/MobileRT/Boxes/AABB.hpp. */
class AABB final {
public:
    vec3 pointMin_{};
    vec3 pointMax_{};

public:
    AABB() noexcept = default;
    explicit AABB(const vec3& pointMin, const vec3& pointMax) noexcept;
    float getSurfaceArea() const noexcept;
};
AABB surroundingBox(const AABB& box1, const AABB& box2) noexcept;
/* AABB.cpp:15-35 slab test of one ray (mrt_box_rays). */
bool intersect(const AABB& box, const Ray& ray) noexcept;

/* app/MobileRT/Intersection.hpp. */
class Intersection final {
public:
    vec3 point_{};
    vec3 normal_{};
    const Material* material_{nullptr};
    float length_{RayLengthMax};
    const void* primitive_{nullptr};
    std::int32_t materialId_{};
    vec3 primitiveColor_{};

public:
    explicit Intersection(float dist) noexcept;
    explicit Intersection(const vec3& intPoint, float dist, const vec3& normal, const void* primitive,
                          std::int32_t materialId, vec3 primitiveColor = vec3{1.0f}) noexcept;
};

/* app/MobileRT/Shapes/Triangle.hpp (same members; construction through
 * mrt_triangle_make, so AB/AC and the face normal round as the GPU's). */
class Triangle final {
public:
    vec3 AC_{};
    vec3 AB_{};
    vec3 pointA_{};
    vec3 normalA_{};
    vec3 normalB_{};
    vec3 normalC_{};
    vec3 colorA_{};
    vec3 colorB_{};
    vec3 colorC_{};
    std::int32_t materialId_{};

public:
    explicit Triangle(const vec3& pointA, const vec3& pointB, const vec3& pointC, std::int32_t materialId) noexcept;
    explicit Triangle(const vec3& pointA, const vec3& pointB, const vec3& pointC, std::int32_t materialId,
                      const vec3& normalA, const vec3& normalB, const vec3& normalC, const vec3& colorA,
                      const vec3& colorB, const vec3& colorC) noexcept;
    explicit Triangle(const vec3& pointA, const vec3& pointB, const vec3& pointC, std::int32_t materialId,
                      const vec3& colorA, const vec3& colorB, const vec3& colorC) noexcept;
    explicit Triangle(const mrt_triangle& t) noexcept;
    bool intersect(Intersection* intersection, const Ray& ray) const noexcept;
    bool intersect(const Ray& ray, float dist) const noexcept;
    void moveTo(float x, float y) noexcept;
    float getZ() const noexcept;
    AABB getAABB() const noexcept;
    bool intersectBox(const AABB& box) const noexcept;
    mrt_triangle descriptor() const noexcept;
};

/* app/MobileRT/Shapes/Sphere.hpp. */
class Sphere final {
private:
    vec3 center_{};
    float sq_radius_{};
    std::int32_t materialId_{};

public:
    explicit Sphere(const vec3& center, float radius, std::int32_t materialId) noexcept;
    bool intersect(Intersection* intersection, const Ray& ray) const noexcept;
    bool intersect(const Ray& ray, float dist) const noexcept;
    void moveTo(float x, float y) noexcept;
    float getZ() const noexcept;
    AABB getAABB() const noexcept;
    bool intersectBox(const AABB& box) const noexcept;
    mrt_sphere descriptor() const noexcept;
};

/* app/MobileRT/Shapes/Plane.hpp. */
class Plane final {
private:
    vec3 normal_{};
    vec3 point_{};
    std::int32_t materialId_{};

public:
    explicit Plane(const vec3& point, const vec3& normal, std::int32_t materialId) noexcept;
    bool intersect(Intersection* intersection, const Ray& ray) const noexcept;
    bool intersect(const Ray& ray, float dist) const noexcept;
    void moveTo(float x, float y) noexcept;
    float getZ() const noexcept;
    AABB getAABB() const noexcept;
    float distance(const vec3& point) const noexcept;
    bool intersectBox(const AABB& box) const noexcept;
    mrt_plane descriptor() const noexcept;
};

/* app/MobileRT/Sampler.hpp.  getSample keeps the reference's shared-counter
 * semantics on the host; descriptor() names the device sampler that renders
 * use for the same role (pixel jitter, Russian roulette, light point). */
class Sampler {
public:
    std::atomic<std::uint32_t> sample_{0};
    const std::uint32_t domainSize_{std::numeric_limits<std::uint32_t>::max()};
    std::uint32_t samples_{std::numeric_limits<std::uint32_t>::max()};

public:
    explicit Sampler() noexcept = default;
    /* Sampler.cpp:9-13: domainSize = (w / (w/16)) * (h / (w/16)). */
    explicit Sampler(std::uint32_t width, std::uint32_t height, std::uint32_t samples) noexcept;
    Sampler(const Sampler&) = delete;
    Sampler& operator=(const Sampler&) = delete;
    virtual ~Sampler() noexcept;
    void resetSampling() noexcept;
    void stopSampling() noexcept;
    virtual float getSample(std::uint32_t sample) noexcept = 0;
    float getSample() noexcept;
    /* pixel: the sampler jitters camera rays (the *_COMPAT counter kinds
     * exist for that role only). */
    virtual mrt_sampler descriptor(bool pixel) const noexcept = 0;
};

/* app/MobileRT/Camera.hpp. */
class Camera {
protected:
    std::atomic<std::uint32_t> block_{0};
    std::uint32_t tileSeed_{0};

protected:
    float degToRad(float deg) const noexcept;
    float radToDeg(float rad) const noexcept;

public:
    vec3 position_{};
    vec3 direction_{};
    vec3 right_{};
    vec3 up_{};

public:
    explicit Camera(const vec3& position, const vec3& lookAt, const vec3& up) noexcept;
    Camera(const Camera& camera) noexcept;
    virtual ~Camera() noexcept;
    virtual Ray generateRay(float u, float v, float deviationU, float deviationV) const noexcept = 0;
    /* Camera.cpp:13-37 tile scheduler over the 256-entry tile table. */
    float getBlock(std::uint32_t sample) noexcept;
    virtual AABB getAABB() const noexcept;
    void resetSampling() noexcept;
    /* Seed of the tile-order shuffle (the reference seeds it from
     * /dev/urandom once per process; 0 keeps the Halton order). */
    void setTileSeed(std::uint32_t seed) noexcept { tileSeed_ = seed; }
    std::uint32_t tileSeed() const noexcept { return tileSeed_; }
    virtual mrt_camera descriptor() const noexcept = 0;
};

/* app/MobileRT/Light.hpp. */
class Light {
public:
    const Material radiance_;

public:
    explicit Light(const Material& radiance) noexcept;
    virtual ~Light() noexcept;
    virtual vec3 getPosition() noexcept = 0;
    virtual void resetSampling() noexcept = 0;
    virtual bool intersect(Intersection* intersection, const Ray& ray) const noexcept = 0;
    virtual mrt_light descriptor() const noexcept = 0;
};

/* app/MobileRT/Scene.hpp. */
class Scene final {
public:
    std::vector<Triangle> triangles_{};
    std::vector<Sphere> spheres_{};
    std::vector<Plane> planes_{};
    std::vector<Material> materials_{};
    std::vector<std::unique_ptr<Light>> lights_{};

public:
    Scene() noexcept = default;
    Scene(Scene&&) noexcept = default;
    Scene& operator=(Scene&&) noexcept = default;
    ~Scene() noexcept;
    /* Scene.cpp:31-37: every light in index order, t < length_. */
    bool traceLights(Intersection* intersection, const Ray& ray) const noexcept;
    void resetSampling() noexcept;
};

class Shader;

namespace detail {
/* Device scene of one primitive category under one accelerator (the
 * accelerator templates' backing store). */
class CategoryScene {
public:
    CategoryScene() noexcept = default;
    CategoryScene(const CategoryScene&) = delete;
    CategoryScene& operator=(const CategoryScene&) = delete;
    CategoryScene(CategoryScene&& o) noexcept;
    CategoryScene& operator=(CategoryScene&& o) noexcept;
    ~CategoryScene() noexcept;
    bool build(int category, const void* prims, std::uint32_t n, int accelerator, const AABB& camera) noexcept;
    /* Closest hit / any hit of one ray; *index = primitive index. */
    bool closest(const Ray& ray, std::uint32_t exclude, float* t, std::uint32_t* index, float* u,
                 float* v) noexcept;
    bool any(const Ray& ray, std::uint32_t exclude, float dist) noexcept;
    bool built() const noexcept { return scene_ != nullptr; }

private:
    mrt_scene* scene_{nullptr};
};

/* Fills an Intersection from a device hit on primitive `prim` (point o + d t,
 * interpolated normal and colour as Triangle.cpp:88-101 etc.). */
void fill(Intersection* is, const Ray& ray, float t, float u, float v, const Triangle& prim) noexcept;
void fill(Intersection* is, const Ray& ray, float t, const Sphere& prim) noexcept;
void fill(Intersection* is, const Ray& ray, float t, const Plane& prim) noexcept;
template <class T> constexpr int category() noexcept;
template <> constexpr int category<Plane>() noexcept { return MRT_CAT_PLANE; }
template <> constexpr int category<Sphere>() noexcept { return MRT_CAT_SPHERE; }
template <> constexpr int category<Triangle>() noexcept { return MRT_CAT_TRIANGLE; }
/* Device id of the primitive a ray leaves (MRT_ID_NONE when not in `prims`). */
template <class T>
std::uint32_t exclude_id(const std::vector<T>& prims, const void* p) noexcept {
    if (p == nullptr || prims.empty()) return MRT_ID_NONE;
    const T* b = prims.data();
    const T* q = static_cast<const T*>(p);
    if (q < b || q >= b + prims.size()) return MRT_ID_NONE;
    return MRT_MAKE_ID(category<T>(), static_cast<std::uint32_t>(q - b));
}
std::vector<mrt_plane> descriptors(const std::vector<Plane>& v);
std::vector<mrt_sphere> descriptors(const std::vector<Sphere>& v);
std::vector<mrt_triangle> descriptors(const std::vector<Triangle>& v);
}  // namespace detail

/* The accelerator templates (Naive.hpp, RegularGrid.hpp, BVH.hpp): the
 * primitives stay on the host in primitives_ (as in the reference, moved in
 * by Shader::initializeAccelerators); the first trace / shadowTrace builds
 * the category's own device scene under the matching structure, and every
 * query is a one-ray device batch. */
template <class T, int KIND>
class DeviceAccelerator {
public:
    std::vector<T> primitives_{};

public:
    DeviceAccelerator() noexcept = default;
    /* camera: the camera box that widens the RegularGrid bounds
     * (Shader.cpp:52-62). */
    explicit DeviceAccelerator(std::vector<T>&& primitives, const AABB& camera = AABB{}) noexcept
        : primitives_{std::move(primitives)}, camera_{camera} {}
    DeviceAccelerator(DeviceAccelerator&&) noexcept = default;
    DeviceAccelerator& operator=(DeviceAccelerator&&) noexcept = default;
    ~DeviceAccelerator() noexcept = default;

    /* Closest primitive with t < intersection->length_. */
    bool trace(Intersection* intersection, const Ray& ray) noexcept {
        if (primitives_.empty() || !ready()) return false;
        float t, u, v;
        std::uint32_t idx;
        if (!dev_.closest(ray, detail::exclude_id(primitives_, ray.primitive_), &t, &idx, &u, &v)) return false;
        if (!(t < intersection->length_)) return false;
        fill_from(intersection, ray, t, u, v, primitives_[idx]);
        return true;
    }
    bool shadowTrace(const Ray& ray, float dist) noexcept {
        if (primitives_.empty() || !ready()) return false;
        return dev_.any(ray, detail::exclude_id(primitives_, ray.primitive_), dist);
    }

private:
    bool ready() noexcept {
        if (!dev_.built()) {
            const auto d = detail::descriptors(primitives_);
            dev_.build(detail::category<T>(), d.data(), static_cast<std::uint32_t>(d.size()), KIND, camera_);
        }
        return dev_.built();
    }
    static void fill_from(Intersection* is, const Ray& r, float t, float u, float v, const Triangle& p) noexcept {
        detail::fill(is, r, t, u, v, p);
    }
    static void fill_from(Intersection* is, const Ray& r, float t, float, float, const Sphere& p) noexcept {
        detail::fill(is, r, t, p);
    }
    static void fill_from(Intersection* is, const Ray& r, float t, float, float, const Plane& p) noexcept {
        detail::fill(is, r, t, p);
    }
    AABB camera_{};
    detail::CategoryScene dev_{};
};

template <class T> using Naive = DeviceAccelerator<T, MRT_NAIVE>;
template <class T> using RegularGrid = DeviceAccelerator<T, MRT_REGULAR_GRID>;
template <class T> using BVH = DeviceAccelerator<T, MRT_BVH>;

/* app/MobileRT/Shader.hpp.  The scene lives on the device from
 * initializeAccelerators on; rayTrace runs the concrete shader's whole
 * recursion there. */
class Shader {
public:
    Scene scene_{};
    Naive<Plane> naivePlanes_{};
    Naive<Sphere> naiveSpheres_{};
    Naive<Triangle> naiveTriangles_{};
    RegularGrid<Plane> regularGridPlanes_{};
    RegularGrid<Sphere> regularGridSpheres_{};
    RegularGrid<Triangle> regularGridTriangles_{};
    ::MobileRT::BVH<Plane> bvhPlanes_{};
    ::MobileRT::BVH<Sphere> bvhSpheres_{};
    ::MobileRT::BVH<Triangle> bvhTriangles_{};
    enum Accelerator { NAIVE = 0, REGULAR_GRID, BVH };

private:
    const Accelerator accelerator_{};
    mrt_scene* device_{nullptr};
    std::vector<mrt_plane> devPlanes_{};
    std::vector<mrt_sphere> devSpheres_{};
    std::vector<mrt_triangle> devTriangles_{};

protected:
    const std::uint32_t samplesLight_{};
    std::int32_t rayDepthMin_{RayDepthMin};
    std::int32_t rayDepthMax_{RayDepthMax};
    mrt_sampler samplerLight_{MRT_SAMPLER_RANDOM, 0.0f, 0x6c696768u, 0u};  /* Shader::getLight draws   */
    mrt_sampler samplerHemi_{MRT_SAMPLER_RANDOM, 0.0f, 0x68656d69u, 0u};   /* hemisphere draws         */

protected:
    virtual bool shade(vec3* rgb, const Intersection& intersection, const Ray& ray) noexcept = 0;
    /* Device shader kind and its Russian-roulette sampler. */
    virtual int kind() const noexcept = 0;
    virtual mrt_sampler samplerRussianRoulette() const noexcept;
    /* Shader-specific descriptor fields (DepthMap's maxPoint). */
    virtual void describeShader(mrt_render_desc* /*desc*/) const noexcept {}

public:
    void initializeAccelerators(Camera* camera) noexcept;

public:
    explicit Shader(Scene&& scene, std::uint32_t samplesLight, Accelerator accelerator) noexcept;
    Shader(const Shader&) = delete;
    Shader& operator=(const Shader&) = delete;
    virtual ~Shader() noexcept;
    bool rayTrace(vec3* rgb, const Ray& ray) noexcept;
    bool shadowTrace(const Ray& ray, float dist) noexcept;
    virtual void resetSampling() noexcept;
    Light& getLight();

    /* Build-specific: the PathTracer / Whitted depth limits (the reference's
     * compile-time RayDepthMin / RayDepthMax) and the light-choice and
     * hemisphere draw streams. */
    void setRayDepth(std::int32_t depthMin, std::int32_t depthMax) noexcept;
    void setDrawSamplers(const mrt_sampler& light, const mrt_sampler& hemisphere) noexcept;
    /* The render descriptor's shader part for a frame of this shader. */
    void describe(mrt_render_desc* desc) const noexcept;
    mrt_scene* device() const noexcept { return device_; }
    Accelerator accelerator() const noexcept { return accelerator_; }

private:
    std::uint32_t excludeId(const void* primitive) const noexcept;
    bool upload(Camera* camera) noexcept;
};

/* app/MobileRT/Renderer.hpp. */
class Renderer final {
public:
    std::unique_ptr<Camera> camera_{};
    std::unique_ptr<Shader> shader_{};

private:
    std::unique_ptr<Sampler> samplerPixel_{};
    std::uint32_t blockSizeX_{};
    std::uint32_t blockSizeY_{};
    std::uint32_t sample_{};
    const std::uint32_t width_{};
    const std::uint32_t height_{};
    const std::uint32_t domainSize_{};
    const std::uint32_t resolution_{};
    const std::uint32_t samplesPixel_{};
    mrt_render_stats stats_{};
    int kernel_{MRT_KERNEL_AUTO};
    std::vector<float> accum_{};
    bool keepAccum_{false};

public:
    explicit Renderer(std::unique_ptr<Shader> shader, std::unique_ptr<Camera> camera,
                      std::unique_ptr<Sampler> samplerPixel, std::uint32_t width, std::uint32_t height,
                      std::uint32_t samplesPixel) noexcept;
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;
    ~Renderer() noexcept = default;
    /* Renderer.cpp:31-59: every sample pass over the 16x16 tile grid, folded
     * into the RGBA8 bitmap with incrementalAvg.  stride is in bytes. */
    void renderFrame(std::uint32_t* bitmap, std::int32_t numThreads, std::uint32_t stride) noexcept;
    void stopRender() noexcept;
    std::uint32_t getSample() const noexcept;

    /* Build-specific: statistics of the last frame, the kernel shape, and an
     * optional FP32 radiance sum per pixel (3 floats, width-major). */
    const mrt_render_stats& lastStats() const noexcept { return stats_; }
    void setKernel(int kernel) noexcept { kernel_ = kernel; }
    void keepAccumulation(bool on) noexcept { keepAccum_ = on; }
    const std::vector<float>& accumulation() const noexcept { return accum_; }
};

}  // namespace MobileRT
