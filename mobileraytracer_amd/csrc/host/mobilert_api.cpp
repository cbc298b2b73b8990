// MobileRT's C++ engine and plugin API (include/mobilert/*.hpp) over the
// C ABI of libmrt_gpu.so.  Compute goes to the GPU through the C ABI; the
// host keeps what the reference keeps on the host around the path: sampler
// counters, the tile scheduler, object construction and hit-record assembly.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <thread>

#include "../../../include/mobilert/Components.hpp"
#include "../../../include/mobilert/c_wrapper.h"

namespace MobileRT {

namespace {

thread_local std::string g_error;

bool ok(int rc, const char* where) noexcept {
    if (rc == MRT_OK) return true;
    g_error = std::string(where) + " failed (" + std::to_string(rc) + "): " + mrt_last_error();
    std::fprintf(stderr, "MobileRT: %s\n", g_error.c_str());
    return false;
}

bool fail(const std::string& msg) noexcept {
    g_error = msg;
    std::fprintf(stderr, "MobileRT: %s\n", msg.c_str());
    return false;
}

// One process-wide context on device 0 for per-ray calls and shader scenes.
// It lives until the process ends (HIP objects are not torn down in static
// destructors).
mrt_ctx* context() noexcept {
    static mrt_ctx* c = [] {
        mrt_ctx* h = nullptr;
        return ok(mrt_ctx_create(0, &h), "mrt_ctx_create") ? h : nullptr;
    }();
    return c;
}

void st3(float* p, const vec3& v) noexcept {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}
vec3 ld3(const float* p) noexcept { return vec3{p[0], p[1], p[2]}; }

mrt_ray to_ray(const Ray& r, float tmax, std::uint32_t exclude) noexcept {
    mrt_ray m{};
    st3(m.org, r.origin_);
    st3(m.dir, r.direction_);
    m.tmax = tmax;
    m.exclude = exclude;
    return m;
}

// glm formulas in the order the device uses: dot = (x x + y y) + z z,
// normalize = v * (1 / sqrt(dot(v, v))).
float dot3(const vec3& a, const vec3& b) noexcept { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
vec3 normalize3(const vec3& v) noexcept { return v * (1.0f / std::sqrt(dot3(v, v))); }

// One closest/any pair test on the device (ray against one primitive).
bool pair_test(int category, const void* prim, const Ray& ray, float tmax, std::uint32_t exclude, mrt_hit* hit,
               bool* any) noexcept {
    mrt_ctx* c = context();
    if (!c) return false;
    const mrt_ray r = to_ray(ray, tmax, exclude);
    mrt_hit h{};
    std::uint8_t a = 0;
    if (!ok(mrt_intersect_pairs(c, category, prim, &r, 1, &h, &a, MRT_MEM_HOST), "mrt_intersect_pairs")) return false;
    if (hit) *hit = h;
    if (any) *any = a != 0;
    return true;
}

mrt_sampler sampler_desc(const Sampler* s, bool pixel, const mrt_sampler& fallback) noexcept {
    return s ? s->descriptor(pixel) : fallback;
}

}  // namespace

const std::string& lastError() noexcept { return g_error; }

// ---------------------------------------------------------------- basics --
Ray::Ray(const vec3& dir, const vec3& origin, std::int32_t depth, const void* primitive) noexcept
    : origin_{origin}, direction_{dir}, depth_{depth}, primitive_{primitive} {}

Material::Material(const vec3& Kd, const vec3& Ks, const vec3& Kt, float refractiveIndice, const vec3& Le) noexcept
    : Le_{Le}, Kd_{Kd}, Ks_{Ks}, Kt_{Kt}, refractiveIndice_{refractiveIndice} {}

bool Material::operator==(const Material& m) const noexcept {
    auto near3 = [](const vec3& a, const vec3& b) {
        return std::fabs(a.x - b.x) < Epsilon && std::fabs(a.y - b.y) < Epsilon && std::fabs(a.z - b.z) < Epsilon;
    };
    return near3(Le_, m.Le_) && near3(Kd_, m.Kd_) && near3(Ks_, m.Ks_) && near3(Kt_, m.Kt_) &&
           std::fabs(refractiveIndice_ - m.refractiveIndice_) < Epsilon;
}

mrt_material Material::descriptor() const noexcept {
    mrt_material d{};
    st3(d.le, Le_);
    st3(d.kd, Kd_);
    st3(d.ks, Ks_);
    st3(d.kt, Kt_);
    d.ior = refractiveIndice_;
    return d;
}

AABB::AABB(const vec3& pointMin, const vec3& pointMax) noexcept : pointMin_{pointMin}, pointMax_{pointMax} {}

// AABB.cpp:49-60.
float AABB::getSurfaceArea() const noexcept {
    const vec3 lengths = pointMax_ - pointMin_;
    const float lx = lengths.x, ly = lengths.y, lz = lengths.z;
    return 2.0f * ((lx * ly) + (ly * lz) + (lz * lx));
}

AABB surroundingBox(const AABB& box1, const AABB& box2) noexcept {
    auto mn = [](float a, float b) { return b < a ? b : a; };
    auto mx = [](float a, float b) { return a < b ? b : a; };
    return AABB{vec3{mn(box1.pointMin_.x, box2.pointMin_.x), mn(box1.pointMin_.y, box2.pointMin_.y),
                     mn(box1.pointMin_.z, box2.pointMin_.z)},
                vec3{mx(box1.pointMax_.x, box2.pointMax_.x), mx(box1.pointMax_.y, box2.pointMax_.y),
                     mx(box1.pointMax_.z, box2.pointMax_.z)}};
}

bool intersect(const AABB& box, const Ray& ray) noexcept {
    mrt_ctx* c = context();
    if (!c) return false;
    float b6[6];
    st3(b6, box.pointMin_);
    st3(b6 + 3, box.pointMax_);
    const mrt_ray r = to_ray(ray, RayLengthMax, MRT_ID_NONE);
    std::uint8_t hit = 0;
    if (!ok(mrt_box_rays(c, b6, &r, 1, &hit, nullptr, MRT_MEM_HOST), "mrt_box_rays")) return false;
    return hit != 0;
}

Intersection::Intersection(float dist) noexcept : length_{dist} {}

Intersection::Intersection(const vec3& intPoint, float dist, const vec3& normal, const void* primitive,
                           std::int32_t materialId, vec3 primitiveColor) noexcept
    : point_{intPoint}, normal_{normal}, length_{dist}, primitive_{primitive}, materialId_{materialId},
      primitiveColor_{primitiveColor} {}

// ------------------------------------------------------------- triangles --
Triangle::Triangle(const mrt_triangle& t) noexcept
    : AC_{ld3(t.ac)}, AB_{ld3(t.ab)}, pointA_{ld3(t.a)}, normalA_{ld3(t.na)}, normalB_{ld3(t.nb)},
      normalC_{ld3(t.nc)}, colorA_{ld3(t.ca)}, colorB_{ld3(t.cb)}, colorC_{ld3(t.cc)}, materialId_{t.material_id} {}

namespace {
mrt_triangle make_tri(const vec3& a, const vec3& b, const vec3& c, std::int32_t mat, const vec3* n,
                      const vec3* col) noexcept {
    float pa[3], pb[3], pc[3], nn[9], cc[9];
    st3(pa, a);
    st3(pb, b);
    st3(pc, c);
    if (n) {
        st3(nn, n[0]);
        st3(nn + 3, n[1]);
        st3(nn + 6, n[2]);
    }
    if (col) {
        st3(cc, col[0]);
        st3(cc + 3, col[1]);
        st3(cc + 6, col[2]);
    }
    mrt_triangle t{};
    ok(mrt_triangle_make(pa, pb, pc, mat, n ? nn : nullptr, col ? cc : nullptr, &t), "mrt_triangle_make");
    return t;
}
}  // namespace

Triangle::Triangle(const vec3& pointA, const vec3& pointB, const vec3& pointC, std::int32_t materialId) noexcept
    : Triangle(make_tri(pointA, pointB, pointC, materialId, nullptr, nullptr)) {}

Triangle::Triangle(const vec3& pointA, const vec3& pointB, const vec3& pointC, std::int32_t materialId,
                   const vec3& normalA, const vec3& normalB, const vec3& normalC, const vec3& colorA,
                   const vec3& colorB, const vec3& colorC) noexcept
    : Triangle([&] {
          const vec3 n[3] = {normalA, normalB, normalC};
          const vec3 c[3] = {colorA, colorB, colorC};
          return make_tri(pointA, pointB, pointC, materialId, n, c);
      }()) {}

Triangle::Triangle(const vec3& pointA, const vec3& pointB, const vec3& pointC, std::int32_t materialId,
                   const vec3& colorA, const vec3& colorB, const vec3& colorC) noexcept
    : Triangle([&] {
          const vec3 c[3] = {colorA, colorB, colorC};
          return make_tri(pointA, pointB, pointC, materialId, nullptr, c);
      }()) {}

mrt_triangle Triangle::descriptor() const noexcept {
    mrt_triangle t{};
    st3(t.ac, AC_);
    st3(t.ab, AB_);
    st3(t.a, pointA_);
    st3(t.na, normalA_);
    st3(t.nb, normalB_);
    st3(t.nc, normalC_);
    st3(t.ca, colorA_);
    st3(t.cb, colorB_);
    st3(t.cc, colorC_);
    t.material_id = materialId_;
    return t;
}

// Triangle.cpp:57-103 (the ray leaving this triangle never hits it).
bool Triangle::intersect(Intersection* intersection, const Ray& ray) const noexcept {
    if (ray.primitive_ == this) return false;
    const mrt_triangle d = descriptor();
    mrt_hit h{};
    if (!pair_test(MRT_CAT_TRIANGLE, &d, ray, intersection->length_, MRT_ID_NONE, &h, nullptr)) return false;
    if (h.id == MRT_ID_NONE) return false;
    detail::fill(intersection, ray, h.t, h.u, h.v, *this);
    return true;
}

// Triangle.cpp:105-136.
bool Triangle::intersect(const Ray& ray, float dist) const noexcept {
    if (ray.primitive_ == this) return false;
    const mrt_triangle d = descriptor();
    bool any = false;
    return pair_test(MRT_CAT_TRIANGLE, &d, ray, dist, MRT_ID_NONE, nullptr, &any) && any;
}

void Triangle::moveTo(float, float) noexcept {}
float Triangle::getZ() const noexcept { return 0.0f; }

AABB Triangle::getAABB() const noexcept {
    const mrt_triangle d = descriptor();
    float b[6] = {};
    ok(mrt_primitive_aabbs(MRT_CAT_TRIANGLE, &d, 1, b), "mrt_primitive_aabbs");
    return AABB{ld3(b), ld3(b + 3)};
}

bool Triangle::intersectBox(const AABB& box) const noexcept {
    const mrt_triangle d = descriptor();
    float mn[3], mx[3];
    st3(mn, box.pointMin_);
    st3(mx, box.pointMax_);
    const int r = mrt_triangle_intersect_box(&d, mn, mx);
    return r > 0;
}

// --------------------------------------------------------------- spheres --
Sphere::Sphere(const vec3& center, float radius, std::int32_t materialId) noexcept
    : center_{center}, sq_radius_{radius * radius}, materialId_{materialId} {}

mrt_sphere Sphere::descriptor() const noexcept {
    mrt_sphere s{};
    st3(s.center, center_);
    s.sq_radius = sq_radius_;
    s.material_id = materialId_;
    return s;
}

bool Sphere::intersect(Intersection* intersection, const Ray& ray) const noexcept {
    const mrt_sphere d = descriptor();
    mrt_hit h{};
    if (!pair_test(MRT_CAT_SPHERE, &d, ray, intersection->length_, MRT_ID_NONE, &h, nullptr)) return false;
    if (h.id == MRT_ID_NONE) return false;
    detail::fill(intersection, ray, h.t, *this);
    return true;
}

bool Sphere::intersect(const Ray& ray, float dist) const noexcept {
    const mrt_sphere d = descriptor();
    bool any = false;
    return pair_test(MRT_CAT_SPHERE, &d, ray, dist, MRT_ID_NONE, nullptr, &any) && any;
}

void Sphere::moveTo(float x, float y) noexcept {
    center_.x = x;
    center_.y = y;
}
float Sphere::getZ() const noexcept { return center_.z; }

AABB Sphere::getAABB() const noexcept {
    const mrt_sphere d = descriptor();
    float b[6] = {};
    ok(mrt_primitive_aabbs(MRT_CAT_SPHERE, &d, 1, b), "mrt_primitive_aabbs");
    return AABB{ld3(b), ld3(b + 3)};
}

bool Sphere::intersectBox(const AABB& box) const noexcept {
    const mrt_sphere d = descriptor();
    float mn[3], mx[3];
    st3(mn, box.pointMin_);
    st3(mx, box.pointMax_);
    return mrt_primitive_intersect_box(MRT_CAT_SPHERE, &d, mn, mx) > 0;
}

// ---------------------------------------------------------------- planes --
Plane::Plane(const vec3& point, const vec3& normal, std::int32_t materialId) noexcept
    : normal_{normalize3(normal)}, point_{point}, materialId_{materialId} {}

mrt_plane Plane::descriptor() const noexcept {
    mrt_plane p{};
    st3(p.normal, normal_);
    st3(p.point, point_);
    p.material_id = materialId_;
    return p;
}

bool Plane::intersect(Intersection* intersection, const Ray& ray) const noexcept {
    if (ray.primitive_ == this) return false;
    const mrt_plane d = descriptor();
    mrt_hit h{};
    if (!pair_test(MRT_CAT_PLANE, &d, ray, intersection->length_, MRT_ID_NONE, &h, nullptr)) return false;
    if (h.id == MRT_ID_NONE) return false;
    detail::fill(intersection, ray, h.t, *this);
    return true;
}

bool Plane::intersect(const Ray& ray, float dist) const noexcept {
    if (ray.primitive_ == this) return false;
    const mrt_plane d = descriptor();
    bool any = false;
    return pair_test(MRT_CAT_PLANE, &d, ray, dist, MRT_ID_NONE, nullptr, &any) && any;
}

void Plane::moveTo(float, float) noexcept {}
float Plane::getZ() const noexcept { return 0.0f; }

AABB Plane::getAABB() const noexcept {
    const mrt_plane d = descriptor();
    float b[6] = {};
    ok(mrt_primitive_aabbs(MRT_CAT_PLANE, &d, 1, b), "mrt_primitive_aabbs");
    return AABB{ld3(b), ld3(b + 3)};
}

// Plane.cpp:106-121.
float Plane::distance(const vec3& point) const noexcept {
    const float d = (normal_.x * -point_.x + normal_.y * -point_.y) + normal_.z * -point_.z;
    const float num = ((normal_.x * point.x + normal_.y * point.y) + normal_.z * point.z) + d;
    const float den = std::sqrt((normal_.x * normal_.x + normal_.y * normal_.y) + normal_.z * normal_.z);
    return num / den;
}

bool Plane::intersectBox(const AABB& box) const noexcept {
    const mrt_plane d = descriptor();
    float mn[3], mx[3];
    st3(mn, box.pointMin_);
    st3(mx, box.pointMax_);
    return mrt_primitive_intersect_box(MRT_CAT_PLANE, &d, mn, mx) > 0;
}

// ----------------------------------------------------------- hit records --
namespace detail {

// Triangle.cpp:88-101: point o + d t, vertex normal and colour interpolated
// with w = 1 - u - v (not renormalised).
void fill(Intersection* is, const Ray& ray, float t, float u, float v, const Triangle& p) noexcept {
    const float w = (1.0f - u) - v;
    is->point_ = ray.origin_ + ray.direction_ * t;
    is->normal_ = (p.normalA_ * w + p.normalB_ * u) + p.normalC_ * v;
    is->primitiveColor_ = (p.colorA_ * w + p.colorB_ * u) + p.colorC_ * v;
    is->length_ = t;
    is->primitive_ = &p;
    is->materialId_ = p.materialId_;
    is->material_ = nullptr;
}

// Sphere.cpp:42-50: the normal points out of the centre; no self exclusion.
void fill(Intersection* is, const Ray& ray, float t, const Sphere& p) noexcept {
    const mrt_sphere d = p.descriptor();
    is->point_ = ray.origin_ + ray.direction_ * t;
    is->normal_ = normalize3(is->point_ - ld3(d.center));
    is->primitiveColor_ = vec3{1.0f};
    is->length_ = t;
    is->primitive_ = nullptr;
    is->materialId_ = d.material_id;
    is->material_ = nullptr;
}

// Plane.cpp:38-44.
void fill(Intersection* is, const Ray& ray, float t, const Plane& p) noexcept {
    const mrt_plane d = p.descriptor();
    is->point_ = ray.origin_ + ray.direction_ * t;
    is->normal_ = ld3(d.normal);
    is->primitiveColor_ = vec3{1.0f};
    is->length_ = t;
    is->primitive_ = &p;
    is->materialId_ = d.material_id;
    is->material_ = nullptr;
}

std::vector<mrt_plane> descriptors(const std::vector<Plane>& v) {
    std::vector<mrt_plane> d;
    d.reserve(v.size());
    for (const Plane& p : v) d.push_back(p.descriptor());
    return d;
}
std::vector<mrt_sphere> descriptors(const std::vector<Sphere>& v) {
    std::vector<mrt_sphere> d;
    d.reserve(v.size());
    for (const Sphere& p : v) d.push_back(p.descriptor());
    return d;
}
std::vector<mrt_triangle> descriptors(const std::vector<Triangle>& v) {
    std::vector<mrt_triangle> d;
    d.reserve(v.size());
    for (const Triangle& p : v) d.push_back(p.descriptor());
    return d;
}

CategoryScene::CategoryScene(CategoryScene&& o) noexcept : scene_{o.scene_} { o.scene_ = nullptr; }
CategoryScene& CategoryScene::operator=(CategoryScene&& o) noexcept {
    if (this != &o) {
        if (scene_) mrt_scene_destroy(scene_);
        scene_ = o.scene_;
        o.scene_ = nullptr;
    }
    return *this;
}
CategoryScene::~CategoryScene() noexcept {
    if (scene_) mrt_scene_destroy(scene_);
}

bool CategoryScene::build(int category, const void* prims, std::uint32_t n, int accelerator,
                          const AABB& camera) noexcept {
    mrt_ctx* c = context();
    if (!c || n == 0) return false;
    mrt_scene_desc d{};
    if (category == MRT_CAT_TRIANGLE) {
        d.triangles = static_cast<const mrt_triangle*>(prims);
        d.n_triangles = n;
    } else if (category == MRT_CAT_SPHERE) {
        d.spheres = static_cast<const mrt_sphere*>(prims);
        d.n_spheres = n;
    } else {
        d.planes = static_cast<const mrt_plane*>(prims);
        d.n_planes = n;
    }
    // Materials are not needed for traversal; indices are checked against
    // the material count, so give every referenced index one.
    std::int32_t maxMat = -1;
    for (std::uint32_t i = 0; i < n; ++i) {
        const std::int32_t m = category == MRT_CAT_TRIANGLE ? static_cast<const mrt_triangle*>(prims)[i].material_id
                               : category == MRT_CAT_SPHERE ? static_cast<const mrt_sphere*>(prims)[i].material_id
                                                            : static_cast<const mrt_plane*>(prims)[i].material_id;
        maxMat = std::max(maxMat, m);
    }
    std::vector<mrt_material> mats(static_cast<std::size_t>(maxMat + 1));
    d.materials = mats.empty() ? nullptr : mats.data();
    d.n_materials = static_cast<std::uint32_t>(mats.size());
    mrt_scene* s = nullptr;
    if (!ok(mrt_scene_create(c, &d, &s), "mrt_scene_create")) return false;
    mrt_accel_params p{};
    st3(p.camera_min, camera.pointMin_);
    st3(p.camera_max, camera.pointMax_);
    p.leaf_size = 2;
    if (!ok(mrt_scene_build(s, accelerator, &p), "mrt_scene_build")) {
        mrt_scene_destroy(s);
        return false;
    }
    scene_ = s;
    return true;
}

bool CategoryScene::closest(const Ray& ray, std::uint32_t exclude, float* t, std::uint32_t* index, float* u,
                            float* v) noexcept {
    const mrt_ray r = to_ray(ray, RayLengthMax, exclude);
    mrt_hit h{};
    if (!ok(mrt_trace_closest(scene_, &r, 1, &h, MRT_MEM_HOST), "mrt_trace_closest")) return false;
    if (h.id == MRT_ID_NONE) return false;
    *t = h.t;
    *index = MRT_ID_INDEX(h.id);
    *u = h.u;
    *v = h.v;
    return true;
}

bool CategoryScene::any(const Ray& ray, std::uint32_t exclude, float dist) noexcept {
    const mrt_ray r = to_ray(ray, dist, exclude);
    std::uint8_t occ = 0;
    return ok(mrt_trace_any(scene_, &r, 1, &occ, MRT_MEM_HOST), "mrt_trace_any") && occ != 0;
}

}  // namespace detail

// --------------------------------------------------------------- sampler --
Sampler::Sampler(std::uint32_t width, std::uint32_t height, std::uint32_t samples) noexcept
    : domainSize_{width >= 16u ? (width / (width / 16u)) * (height / (width / 16u)) : 0u}, samples_{samples} {}
Sampler::~Sampler() noexcept = default;
void Sampler::resetSampling() noexcept { sample_ = 0; }
void Sampler::stopSampling() noexcept { samples_ = 0; }
float Sampler::getSample() noexcept { return getSample(0); }

// ---------------------------------------------------------------- camera --
Camera::Camera(const vec3& position, const vec3& lookAt, const vec3& up) noexcept : position_{position} {
    float p[3], l[3], u[3];
    st3(p, position);
    st3(l, lookAt);
    st3(u, up);
    mrt_camera c{};
    ok(mrt_camera_make(MRT_CAMERA_PERSPECTIVE, p, l, u, 45.0f, 45.0f, &c), "mrt_camera_make");
    direction_ = ld3(c.direction);
    right_ = ld3(c.right);
    up_ = ld3(c.up);
}

Camera::Camera(const Camera& camera) noexcept
    : tileSeed_{camera.tileSeed_}, position_{camera.position_}, direction_{camera.direction_}, right_{camera.right_},
      up_{camera.up_} {
    block_.store(camera.block_.load());
}
Camera::~Camera() noexcept = default;

float Camera::degToRad(float deg) const noexcept { return (deg * 3.14159265358979323846f) / 180.0f; }
float Camera::radToDeg(float rad) const noexcept { return (rad / 3.14159265358979323846f) * 180.0f; }

// Camera.cpp:28-37 over this camera's tile table (mrt_tile_values).
float Camera::getBlock(std::uint32_t sample) noexcept {
    static std::mutex mu;
    static std::uint32_t seed = 0xFFFFFFFFu;
    static float values[256];
    {
        std::lock_guard<std::mutex> lock(mu);
        if (seed != tileSeed_) {
            ok(mrt_tile_values(tileSeed_, values), "mrt_tile_values");
            seed = tileSeed_;
        }
    }
    const std::uint32_t current = block_.fetch_add(1, std::memory_order_relaxed) - 256u * sample;
    if (current >= 256u) {
        block_.fetch_sub(1, std::memory_order_relaxed);
        return 1.0f;
    }
    return values[current];
}

AABB Camera::getAABB() const noexcept { return AABB{position_, position_}; }
void Camera::resetSampling() noexcept { block_ = 0; }

// ----------------------------------------------------------------- light --
Light::Light(const Material& radiance) noexcept : radiance_{radiance} {}
Light::~Light() noexcept = default;

// ----------------------------------------------------------------- scene --
Scene::~Scene() noexcept = default;

// Scene.cpp:31-37: every light in index order, keeping a hit only when it is
// nearer than the current one.
bool Scene::traceLights(Intersection* intersection, const Ray& ray) const noexcept {
    bool res = false;
    for (const auto& light : lights_) {
        if (light->intersect(intersection, ray)) res = true;
    }
    return res;
}

void Scene::resetSampling() noexcept {
    for (auto& light : lights_) light->resetSampling();
}

// ---------------------------------------------------------------- shader --
Shader::Shader(Scene&& scene, std::uint32_t samplesLight, Accelerator accelerator) noexcept
    : scene_{std::move(scene)}, accelerator_{accelerator}, samplesLight_{samplesLight} {}

Shader::~Shader() noexcept {
    if (device_) mrt_scene_destroy(device_);
}

mrt_sampler Shader::samplerRussianRoulette() const noexcept {
    return mrt_sampler{MRT_SAMPLER_CONSTANT, 0.5f, 0u, 0u};
}

void Shader::setRayDepth(std::int32_t depthMin, std::int32_t depthMax) noexcept {
    rayDepthMin_ = depthMin;
    rayDepthMax_ = depthMax;
}

void Shader::setDrawSamplers(const mrt_sampler& light, const mrt_sampler& hemisphere) noexcept {
    samplerLight_ = light;
    samplerHemi_ = hemisphere;
}

bool Shader::upload(Camera* camera) noexcept {
    mrt_ctx* c = context();
    if (!c) return false;
    devPlanes_ = detail::descriptors(scene_.planes_);
    devSpheres_ = detail::descriptors(scene_.spheres_);
    devTriangles_ = detail::descriptors(scene_.triangles_);
    std::vector<mrt_material> mats;
    for (const Material& m : scene_.materials_) mats.push_back(m.descriptor());
    std::vector<mrt_light> lights;
    for (const auto& l : scene_.lights_) lights.push_back(l->descriptor());
    mrt_scene_desc d{};
    d.triangles = devTriangles_.empty() ? nullptr : devTriangles_.data();
    d.n_triangles = static_cast<std::uint32_t>(devTriangles_.size());
    d.spheres = devSpheres_.empty() ? nullptr : devSpheres_.data();
    d.n_spheres = static_cast<std::uint32_t>(devSpheres_.size());
    d.planes = devPlanes_.empty() ? nullptr : devPlanes_.data();
    d.n_planes = static_cast<std::uint32_t>(devPlanes_.size());
    d.materials = mats.empty() ? nullptr : mats.data();
    d.n_materials = static_cast<std::uint32_t>(mats.size());
    d.lights = lights.empty() ? nullptr : lights.data();
    d.n_lights = static_cast<std::uint32_t>(lights.size());
    if (device_) {
        mrt_scene_destroy(device_);
        device_ = nullptr;
    }
    if (!ok(mrt_scene_create(c, &d, &device_), "mrt_scene_create")) return false;
    mrt_accel_params p{};
    const AABB cam = camera ? camera->getAABB() : AABB{};
    st3(p.camera_min, cam.pointMin_);
    st3(p.camera_max, cam.pointMax_);
    p.leaf_size = 2;
    return ok(mrt_scene_build(device_, static_cast<int>(accelerator_), &p), "mrt_scene_build");
}

// Shader.cpp:18-122: the primitives move from scene_ into the chosen
// accelerators (scene_.triangles_ etc. are empty afterwards); the device
// scene holds all of them plus the lights.
void Shader::initializeAccelerators(Camera* camera) noexcept {
    if (!upload(camera)) return;
    const AABB cam = camera ? camera->getAABB() : AABB{};
    switch (accelerator_) {
        case NAIVE:
            naivePlanes_ = Naive<Plane>{std::move(scene_.planes_), cam};
            naiveSpheres_ = Naive<Sphere>{std::move(scene_.spheres_), cam};
            naiveTriangles_ = Naive<Triangle>{std::move(scene_.triangles_), cam};
            break;
        case REGULAR_GRID:
            regularGridPlanes_ = RegularGrid<Plane>{std::move(scene_.planes_), cam};
            regularGridSpheres_ = RegularGrid<Sphere>{std::move(scene_.spheres_), cam};
            regularGridTriangles_ = RegularGrid<Triangle>{std::move(scene_.triangles_), cam};
            break;
        case BVH:
            bvhPlanes_ = MobileRT::BVH<Plane>{std::move(scene_.planes_), cam};
            bvhSpheres_ = MobileRT::BVH<Sphere>{std::move(scene_.spheres_), cam};
            bvhTriangles_ = MobileRT::BVH<Triangle>{std::move(scene_.triangles_), cam};
            break;
    }
    scene_.planes_.clear();
    scene_.spheres_.clear();
    scene_.triangles_.clear();
}

std::uint32_t Shader::excludeId(const void* primitive) const noexcept {
    if (primitive == nullptr) return MRT_ID_NONE;
    std::uint32_t id = MRT_ID_NONE;
    switch (accelerator_) {
        case NAIVE:
            if ((id = detail::exclude_id(naiveTriangles_.primitives_, primitive)) != MRT_ID_NONE) return id;
            if ((id = detail::exclude_id(naivePlanes_.primitives_, primitive)) != MRT_ID_NONE) return id;
            break;
        case REGULAR_GRID:
            if ((id = detail::exclude_id(regularGridTriangles_.primitives_, primitive)) != MRT_ID_NONE) return id;
            if ((id = detail::exclude_id(regularGridPlanes_.primitives_, primitive)) != MRT_ID_NONE) return id;
            break;
        case BVH:
            if ((id = detail::exclude_id(bvhTriangles_.primitives_, primitive)) != MRT_ID_NONE) return id;
            if ((id = detail::exclude_id(bvhPlanes_.primitives_, primitive)) != MRT_ID_NONE) return id;
            break;
    }
    for (std::size_t i = 0; i < scene_.lights_.size(); ++i) {
        const auto* al = dynamic_cast<const Components::AreaLight*>(scene_.lights_[i].get());
        if (al && primitive == &al->triangle()) return MRT_MAKE_ID(MRT_CAT_LIGHT, static_cast<std::uint32_t>(i));
    }
    return MRT_ID_NONE;
}

void Shader::describe(mrt_render_desc* d) const noexcept {
    d->shader = kind();
    d->samples_light = samplesLight_;
    d->ray_depth_min = rayDepthMin_;
    d->ray_depth_max = rayDepthMax_;
    d->sampler_rr = samplerRussianRoulette();
    d->sampler_light = samplerLight_;
    d->sampler_hemi = samplerHemi_;
    describeShader(d);
}

// Shader.cpp:124-160 and the concrete shade: the ray's whole recursion on
// the device (depth = ray.depth_).
bool Shader::rayTrace(vec3* rgb, const Ray& ray) noexcept {
    if (!device_) return fail("rayTrace before initializeAccelerators");
    mrt_render_desc d{};
    d.width = 16;
    d.height = 16;
    d.samples_pixel = 1;
    d.camera.kind = MRT_CAMERA_PERSPECTIVE;
    describe(&d);
    const mrt_ray r = to_ray(ray, RayLengthMax, excludeId(ray.primitive_));
    const std::int32_t depth = ray.depth_;
    float out[3] = {};
    std::uint8_t flag = 0;
    if (!ok(mrt_shade_rays(device_, &d, &r, &depth, 1, 0, out, &flag, MRT_MEM_HOST), "mrt_shade_rays")) return false;
    *rgb = ld3(out);
    return flag != 0;
}

// Shader.cpp:162-187.
bool Shader::shadowTrace(const Ray& ray, float dist) noexcept {
    if (!device_) return fail("shadowTrace before initializeAccelerators");
    const mrt_ray r = to_ray(ray, dist, excludeId(ray.primitive_));
    std::uint8_t occ = 0;
    return ok(mrt_trace_any(device_, &r, 1, &occ, MRT_MEM_HOST), "mrt_trace_any") && occ != 0;
}

void Shader::resetSampling() noexcept { scene_.resetSampling(); }

// Shader.cpp:225-236: a uniform light index (the draw comes from this
// shader's light-choice stream).
Light& Shader::getLight() {
    static std::atomic<std::uint32_t> counter{0};
    const std::uint32_t n = static_cast<std::uint32_t>(scene_.lights_.size());
    std::uint32_t h = samplerLight_.seed ^ (counter.fetch_add(1) * 0x9E3779B9u);
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    const float r = static_cast<float>(h >> 8) * (1.0f / 16777216.0f);
    const std::uint32_t idx = std::min(n - 1u, static_cast<std::uint32_t>(std::floor(r * n * 0.99999f)));
    return *scene_.lights_[idx];
}

// -------------------------------------------------------------- renderer --
Renderer::Renderer(std::unique_ptr<Shader> shader, std::unique_ptr<Camera> camera,
                   std::unique_ptr<Sampler> samplerPixel, std::uint32_t width, std::uint32_t height,
                   std::uint32_t samplesPixel) noexcept
    : camera_{std::move(camera)}, shader_{std::move(shader)}, samplerPixel_{std::move(samplerPixel)},
      blockSizeX_{width / 16u}, blockSizeY_{height / 16u}, sample_{0}, width_{width}, height_{height},
      domainSize_{blockSizeX_ && blockSizeY_ ? (width / blockSizeX_) * (height / blockSizeY_) : 0u},
      resolution_{width * height}, samplesPixel_{samplesPixel} {
    shader_->initializeAccelerators(camera_.get());
}

void Renderer::renderFrame(std::uint32_t* bitmap, std::int32_t /*numThreads*/, std::uint32_t stride) noexcept {
    if (!shader_ || !shader_->device()) {
        fail("renderFrame: the shader has no device scene");
        return;
    }
    samplerPixel_->resetSampling();
    shader_->resetSampling();
    camera_->resetSampling();
    mrt_render_desc d{};
    d.width = width_;
    d.height = height_;
    d.samples_pixel = samplesPixel_;
    d.camera = camera_->descriptor();
    shader_->describe(&d);
    mrt_sampler pix = sampler_desc(samplerPixel_.get(), true, mrt_sampler{MRT_SAMPLER_CONSTANT, 0.5f, 0u, 0u});
    if ((pix.kind == MRT_SAMPLER_HALTON_COMPAT || pix.kind == MRT_SAMPLER_STRATIFIED_COMPAT) && pix.domain_size == 0)
        pix.domain_size = domainSize_;
    d.sampler_pixel = pix;
    ok(mrt_tile_values(camera_->tileSeed(), d.tile_values), "mrt_tile_values");
    d.kernel = kernel_;
    if (keepAccum_) accum_.assign(static_cast<std::size_t>(width_) * height_ * 3u, 0.0f);
    std::memset(&stats_, 0, sizeof(stats_));
    const int rc = mrt_render(shader_->device(), &d, bitmap, stride, keepAccum_ ? accum_.data() : nullptr,
                              MRT_MEM_HOST, &stats_);
    if (rc == MRT_E_STOPPED) return;
    if (ok(rc, "mrt_render")) sample_ = stats_.samples_done;
}

void Renderer::stopRender() noexcept {
    blockSizeX_ = 0;
    blockSizeY_ = 0;
    samplerPixel_->stopSampling();
    if (shader_ && shader_->device()) mrt_render_stop(shader_->device());
}

std::uint32_t Renderer::getSample() const noexcept {
    std::uint32_t s = sample_;
    if (shader_ && shader_->device()) mrt_render_progress(shader_->device(), &s);
    return s;
}

}  // namespace MobileRT

// ================================================================ plugins ==
namespace Components {

using MobileRT::vec3;

namespace {
// The seeded counter hash of the deterministic contract (one stream per
// seed): u01(fmix32(seed ^ golden * counter)).
float hash_u01(std::uint32_t seed, std::uint32_t counter) noexcept {
    std::uint32_t h = seed ^ (counter * 0x9E3779B9u);
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return static_cast<float>(h >> 8) * (1.0f / 16777216.0f);
}

// Utils.cpp:36-45 radical inverse.
float halton(std::uint32_t index, std::uint32_t base) noexcept {
    float f = 1.0f, result = 0.0f;
    while (index > 0) {
        f /= static_cast<float>(base);
        result += f * static_cast<float>(index % base);
        index /= base;
    }
    return result;
}

std::atomic<std::uint32_t> g_seed{0x5EEDu};
}  // namespace

// ------------------------------------------------------------- shaders --
Whitted::Whitted(MobileRT::Scene&& scene, std::uint32_t samplesLight, Accelerator accelerator) noexcept
    : Shader(std::move(scene), samplesLight, accelerator) {}

// The device shades the intersection the ray produces: tracing the same ray
// again yields this intersection, so shade() is rayTrace() of that ray.
bool Whitted::shade(vec3* rgb, const MobileRT::Intersection&, const MobileRT::Ray& ray) noexcept {
    return rayTrace(rgb, ray);
}

NoShadows::NoShadows(MobileRT::Scene&& scene, std::uint32_t samplesLight, Accelerator accelerator) noexcept
    : Shader(std::move(scene), samplesLight, accelerator) {}

bool NoShadows::shade(vec3* rgb, const MobileRT::Intersection&, const MobileRT::Ray& ray) noexcept {
    return rayTrace(rgb, ray);
}

// DepthMap.cpp:12-16: no light samples.
DepthMap::DepthMap(MobileRT::Scene&& scene, const vec3& maxPoint, Accelerator accelerator) noexcept
    : Shader(std::move(scene), 0, accelerator), maxPoint_{maxPoint} {}

bool DepthMap::shade(vec3* rgb, const MobileRT::Intersection&, const MobileRT::Ray& ray) noexcept {
    return rayTrace(rgb, ray);
}

void DepthMap::describeShader(mrt_render_desc* d) const noexcept {
    d->depth_max_point[0] = maxPoint_.x;
    d->depth_max_point[1] = maxPoint_.y;
    d->depth_max_point[2] = maxPoint_.z;
}

// DiffuseMaterial.cpp:11-14: no light samples.
DiffuseMaterial::DiffuseMaterial(MobileRT::Scene&& scene, Accelerator accelerator) noexcept
    : Shader(std::move(scene), 0, accelerator) {}

bool DiffuseMaterial::shade(vec3* rgb, const MobileRT::Intersection&, const MobileRT::Ray& ray) noexcept {
    return rayTrace(rgb, ray);
}

PathTracer::PathTracer(MobileRT::Scene&& scene, std::unique_ptr<MobileRT::Sampler> samplerRussianRoulette,
                       std::uint32_t samplesLight, Accelerator accelerator) noexcept
    : Shader(std::move(scene), samplesLight, accelerator), samplerRussianRoulette_{std::move(samplerRussianRoulette)} {}

bool PathTracer::shade(vec3* rgb, const MobileRT::Intersection&, const MobileRT::Ray& ray) noexcept {
    return rayTrace(rgb, ray);
}

mrt_sampler PathTracer::samplerRussianRoulette() const noexcept {
    return samplerRussianRoulette_ ? samplerRussianRoulette_->descriptor(false)
                                   : mrt_sampler{MRT_SAMPLER_CONSTANT, 0.5f, 0u, 0u};
}

void PathTracer::resetSampling() noexcept {
    Shader::resetSampling();
    if (samplerRussianRoulette_) samplerRussianRoulette_->resetSampling();
}

// ------------------------------------------------------------ samplers --
Constant::Constant(float value) noexcept : value_{value} {}
float Constant::getSample(std::uint32_t) noexcept { return value_; }
mrt_sampler Constant::descriptor(bool) const noexcept { return mrt_sampler{MRT_SAMPLER_CONSTANT, value_, 0u, 0u}; }

HaltonSeq::HaltonSeq(std::uint32_t width, std::uint32_t height, std::uint32_t samples) noexcept
    : Sampler(width, height, samples) {}

// HaltonSeq.cpp:14-22.
float HaltonSeq::getSample(std::uint32_t sample) noexcept {
    const std::uint32_t current = sample_.fetch_add(1, std::memory_order_relaxed);
    if (current >= domainSize_ * (sample + 1u)) {
        sample_.fetch_sub(1, std::memory_order_relaxed);
        return 1.0f;
    }
    return halton(current - sample * domainSize_, 2u);
}

mrt_sampler HaltonSeq::descriptor(bool pixel) const noexcept {
    if (pixel && !perPixel_) return mrt_sampler{MRT_SAMPLER_HALTON_COMPAT, 0.0f, 0u, domainSize_};
    return mrt_sampler{MRT_SAMPLER_HALTON, 0.0f, seed_, 0u};
}

Stratified::Stratified(std::uint32_t width, std::uint32_t height, std::uint32_t samples) noexcept
    : Sampler(width, height, samples) {}

// Stratified.cpp:14-23.
float Stratified::getSample(std::uint32_t sample) noexcept {
    const std::uint32_t current = sample_.fetch_add(1, std::memory_order_relaxed);
    if (current >= domainSize_ * (sample + 1u)) {
        sample_.fetch_sub(1, std::memory_order_relaxed);
        return 1.0f;
    }
    return static_cast<float>(current - sample * domainSize_) / static_cast<float>(domainSize_);
}

mrt_sampler Stratified::descriptor(bool pixel) const noexcept {
    if (pixel && !perPixel_) return mrt_sampler{MRT_SAMPLER_STRATIFIED_COMPAT, 0.0f, 0u, domainSize_};
    return mrt_sampler{MRT_SAMPLER_STRATIFIED, 0.0f, seed_, 0u};
}

MersenneTwister::MersenneTwister() noexcept : seed_{g_seed.fetch_add(0x2545F491u)} {}
MersenneTwister::MersenneTwister(std::uint32_t seed) noexcept : seed_{seed} {}

float MersenneTwister::getSample(std::uint32_t) noexcept {
    return hash_u01(seed_, sample_.fetch_add(1, std::memory_order_relaxed));
}

mrt_sampler MersenneTwister::descriptor(bool) const noexcept {
    return mrt_sampler{MRT_SAMPLER_RANDOM, 0.0f, seed_, 0u};
}

// ------------------------------------------------------------- cameras --
namespace {
mrt_camera camera_desc(int kind, const MobileRT::Camera& c, float a, float b) noexcept {
    mrt_camera d{};
    d.kind = kind;
    MobileRT::vec3 v = c.position_;
    d.position[0] = v.x, d.position[1] = v.y, d.position[2] = v.z;
    v = c.direction_;
    d.direction[0] = v.x, d.direction[1] = v.y, d.direction[2] = v.z;
    v = c.right_;
    d.right[0] = v.x, d.right[1] = v.y, d.right[2] = v.z;
    v = c.up_;
    d.up[0] = v.x, d.up[1] = v.y, d.up[2] = v.z;
    d.fov_h = a;
    d.fov_v = b;
    return d;
}

MobileRT::Ray device_camera_ray(const mrt_camera& cam, float u, float v, float du, float dv) noexcept {
    float uvdd[4] = {u, v, du, dv};
    mrt_ray r{};
    mrt_ctx* c = MobileRT::context();
    if (!c || !MobileRT::ok(mrt_camera_rays(c, &cam, uvdd, 1, &r, MRT_MEM_HOST), "mrt_camera_rays"))
        return MobileRT::Ray{vec3{}, vec3{}, 1};
    return MobileRT::Ray{vec3{r.dir[0], r.dir[1], r.dir[2]}, vec3{r.org[0], r.org[1], r.org[2]}, 1};
}
}  // namespace

Perspective::Perspective(const vec3& position, const vec3& lookAt, const vec3& up, float hFov, float vFov) noexcept
    : Camera(position, lookAt, up), hFov_{degToRad(hFov)}, vFov_{degToRad(vFov)} {}

mrt_camera Perspective::descriptor() const noexcept {
    return camera_desc(MRT_CAMERA_PERSPECTIVE, *this, hFov_, vFov_);
}

// Perspective.cpp:24-37.
MobileRT::Ray Perspective::generateRay(float u, float v, float deviationU, float deviationV) const noexcept {
    return device_camera_ray(descriptor(), u, v, deviationU, deviationV);
}

float Perspective::getHFov() const noexcept { return radToDeg(hFov_); }
float Perspective::getVFov() const noexcept { return radToDeg(vFov_); }

Orthographic::Orthographic(const vec3& position, const vec3& lookAt, const vec3& up, float sizeH, float sizeV) noexcept
    : Camera(position, lookAt, up), sizeH_{sizeH / 2.0f}, sizeV_{sizeV / 2.0f} {}

mrt_camera Orthographic::descriptor() const noexcept {
    return camera_desc(MRT_CAMERA_ORTHOGRAPHIC, *this, sizeH_, sizeV_);
}

// Orthographic.cpp:23-31.
MobileRT::Ray Orthographic::generateRay(float u, float v, float deviationU, float deviationV) const noexcept {
    return device_camera_ray(descriptor(), u, v, deviationU, deviationV);
}

// Orthographic.cpp:33-45 (the view rectangle widened by half a unit).
MobileRT::AABB Orthographic::getAABB() const noexcept {
    const mrt_camera d = descriptor();
    float b[6] = {};
    MobileRT::ok(mrt_camera_aabb(&d, b), "mrt_camera_aabb");
    return MobileRT::AABB{MobileRT::ld3(b), MobileRT::ld3(b + 3)};
}

float Orthographic::getSizeH() const noexcept { return sizeH_; }
float Orthographic::getSizeV() const noexcept { return sizeV_; }

// -------------------------------------------------------------- lights --
AreaLight::AreaLight(const MobileRT::Material& radiance, std::unique_ptr<MobileRT::Sampler> samplerPointLight,
                     const vec3& pointA, const vec3& pointB, const vec3& pointC) noexcept
    : Light(radiance), triangle_{pointA, pointB, pointC, -1}, samplerPointLight_{std::move(samplerPointLight)} {}

AreaLight::AreaLight(const MobileRT::Material& radiance, std::unique_ptr<MobileRT::Sampler> samplerPointLight,
                     const MobileRT::Triangle& triangle) noexcept
    : Light(radiance), triangle_{triangle}, samplerPointLight_{std::move(samplerPointLight)} {}

// AreaLight.cpp:22-31.
vec3 AreaLight::getPosition() noexcept {
    float R = samplerPointLight_->getSample();
    float S = samplerPointLight_->getSample();
    if (R + S >= 1.0f) {
        R = 1.0f - R;
        S = 1.0f - S;
    }
    return (triangle_.pointA_ + triangle_.AB_ * R) + triangle_.AC_ * S;
}

void AreaLight::resetSampling() noexcept { samplerPointLight_->resetSampling(); }

// AreaLight.cpp:37-43.
bool AreaLight::intersect(MobileRT::Intersection* intersection, const MobileRT::Ray& ray) const noexcept {
    const bool hit = triangle_.intersect(intersection, ray);
    if (hit) intersection->material_ = &radiance_;
    return hit;
}

mrt_light AreaLight::descriptor() const noexcept {
    mrt_light l{};
    l.kind = MRT_LIGHT_AREA;
    l.radiance = radiance_.descriptor();
    l.triangle = triangle_.descriptor();
    l.sampler = samplerPointLight_ ? samplerPointLight_->descriptor(false)
                                   : mrt_sampler{MRT_SAMPLER_CONSTANT, 0.5f, 0u, 0u};
    return l;
}

PointLight::PointLight(const MobileRT::Material& radiance, const vec3& position) noexcept
    : Light(radiance), position_{position} {}
vec3 PointLight::getPosition() noexcept { return position_; }
void PointLight::resetSampling() noexcept {}
bool PointLight::intersect(MobileRT::Intersection*, const MobileRT::Ray&) const noexcept { return false; }

mrt_light PointLight::descriptor() const noexcept {
    mrt_light l{};
    l.kind = MRT_LIGHT_POINT;
    l.radiance = radiance_.descriptor();
    l.position[0] = position_.x;
    l.position[1] = position_.y;
    l.position[2] = position_.z;
    l.triangle.material_id = -1;
    return l;
}

}  // namespace Components

// ------------------------------------------------------------ OBJ loader --
namespace Components {

OBJLoader::OBJLoader(std::string obj, std::string materials) noexcept
    : objFilePath_{std::move(obj)}, mtlFilePath_{std::move(materials)} {}

OBJLoader::~OBJLoader() noexcept {
    if (obj_) mrt_obj_destroy(obj_);
}

// OBJLoader.cpp:23-64.
std::int32_t OBJLoader::process() noexcept {
    if (obj_) {
        mrt_obj_destroy(obj_);
        obj_ = nullptr;
    }
    if (mrt_obj_load(objFilePath_.c_str(), mtlFilePath_.empty() ? nullptr : mtlFilePath_.c_str(), &obj_) != MRT_OK) {
        std::fprintf(stderr, "OBJLoader: %s\n", mrt_last_error());
        isProcessed_ = false;
        numberTriangles_ = -1;
        return numberTriangles_;
    }
    std::uint32_t nt = 0, nl = 0;
    mrt_obj_counts(obj_, &nt, nullptr, &nl, nullptr);
    isProcessed_ = true;
    numberTriangles_ = static_cast<std::int32_t>(nt + nl);
    return numberTriangles_;
}

// OBJLoader.cpp:66-250.
bool OBJLoader::fillScene(MobileRT::Scene* scene, std::function<std::unique_ptr<MobileRT::Sampler>()> lambda) noexcept {
    if (!isProcessed_ || !obj_ || !scene) return false;
    std::uint32_t nt = 0, nm = 0, nl = 0;
    mrt_obj_counts(obj_, &nt, &nm, &nl, nullptr);
    std::vector<mrt_triangle> tris(nt);
    std::vector<mrt_material> mats(nm);
    std::vector<mrt_light> lights(nl);
    if (mrt_obj_fill(obj_, tris.data(), mats.data(), lights.data(), nullptr) != MRT_OK) return false;
    auto material_of = [](const mrt_material& m) {
        return MobileRT::Material{vec3{m.kd[0], m.kd[1], m.kd[2]}, vec3{m.ks[0], m.ks[1], m.ks[2]},
                                  vec3{m.kt[0], m.kt[1], m.kt[2]}, m.ior, vec3{m.le[0], m.le[1], m.le[2]}};
    };
    // The loader's materials against the scene's (first equal one, else
    // appended), as the reference's std::find over scene->materials_.
    std::vector<std::int32_t> index(nm);
    for (std::uint32_t k = 0; k < nm; ++k) {
        const MobileRT::Material m = material_of(mats[k]);
        const auto it = std::find(scene->materials_.begin(), scene->materials_.end(), m);
        if (it != scene->materials_.end()) {
            index[k] = static_cast<std::int32_t>(it - scene->materials_.begin());
        } else {
            index[k] = static_cast<std::int32_t>(scene->materials_.size());
            scene->materials_.push_back(m);
        }
    }
    scene->triangles_.reserve(scene->triangles_.size() + nt);
    for (mrt_triangle t : tris) {
        if (t.material_id >= 0) t.material_id = index[static_cast<std::size_t>(t.material_id)];
        scene->triangles_.emplace_back(t);
    }
    for (const mrt_light& l : lights)
        scene->lights_.emplace_back(
            std::make_unique<AreaLight>(material_of(l.radiance), lambda ? lambda() : std::make_unique<MersenneTwister>(),
                                        MobileRT::Triangle{l.triangle}));
    return true;
}

}  // namespace Components
