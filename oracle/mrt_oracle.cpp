// ORACLE — TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or
// calls this code; tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg use it as the checker.
//
// Scene, shaders, camera and renderer of the CPU restatement (see
// mrt_oracle.hpp for the geometry and accelerators), plus the extern "C"
// surface the Python tests bind with ctypes.
#include "mrt_oracle.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <mutex>

namespace orc {

// ------------------------------------------------------------------- Light --
template <class R>
struct Light {
    int32_t kind = MRT_LIGHT_AREA;
    Material<R> radiance;
    Triangle<R> tri;
    V3<R> position;
    mrt_sampler sampler{};
};

template <class R>
Material<R> mat_from(const mrt_material& m) {
    Material<R> r;
    r.Le = from3<R>(m.le);
    r.Kd = from3<R>(m.kd);
    r.Ks = from3<R>(m.ks);
    r.Kt = from3<R>(m.kt);
    r.ior = static_cast<R>(m.ior);
    return r;
}

// Per-draw context of the deterministic sampling contract.
struct DrawCtx {
    uint32_t pixel = 0;
    uint32_t sample = 0;
    uint32_t branch = 0;
};

template <class R>
struct Counters {
    std::atomic<uint64_t> camera{0}, extension{0}, shadow{0}, specular{0};
};

// ------------------------------------------------------------------- Scene --
template <class R>
struct Scene {
    std::vector<Triangle<R>> tris;
    std::vector<Sphere<R>> sphs;
    std::vector<Plane<R>> plas;
    std::vector<Material<R>> mats;
    std::vector<Light<R>> lights;
    int accel = MRT_NAIVE;
    BVH<R, Plane<R>> bvhP;
    BVH<R, Sphere<R>> bvhS;
    BVH<R, Triangle<R>> bvhT;
    RegularGrid<R, Plane<R>> gridP;
    RegularGrid<R, Sphere<R>> gridS;
    RegularGrid<R, Triangle<R>> gridT;
    BVH<R, Triangle<R>> lightBvh;  // reference-built tree over the area lights (ordered accounting only)
    Naive<R, Plane<R>> naiveP;
    Naive<R, Sphere<R>> naiveS;
    Naive<R, Triangle<R>> naiveT;

    // Shader::initializeAccelerators, Shader.cpp:31-122.  The primitive
    // vectors move into the accelerators; lights and materials stay.
    void build(int kind, const V3<R>& camMin, const V3<R>& camMax, uint32_t leaf) {
        accel = kind;
        if (kind == MRT_NAIVE) {
            naiveP.prims = std::move(plas);
            naiveS.prims = std::move(sphs);
            naiveT.prims = std::move(tris);
        } else if (kind == MRT_REGULAR_GRID) {
            auto bounds = [&](auto& v) {
                V3<R> mn{static_cast<R>(kRayLengthMax), static_cast<R>(kRayLengthMax), static_cast<R>(kRayLengthMax)};
                V3<R> mx{static_cast<R>(-kRayLengthMax), static_cast<R>(-kRayLengthMax), static_cast<R>(-kRayLengthMax)};
                for (const auto& p : v) {  // Scene::getBounds / AABBbounds, Scene.cpp:39-48
                    const AABB<R> b = p.aabb();
                    mn = vmin(b.mn, mn);
                    mx = vmax(b.mx, mx);
                }
                mn = vmin(camMin, mn);
                mx = vmax(camMax, mx);
                return AABB<R>{mn - static_cast<R>(kEpsilon), mx + static_cast<R>(kEpsilon)};
            };
            const AABB<R> bP = bounds(plas), bS = bounds(sphs), bT = bounds(tris);
            const int32_t gP = grid_size_for(static_cast<uint32_t>(plas.size()));
            const int32_t gS = grid_size_for(static_cast<uint32_t>(sphs.size()));
            const int32_t gT = grid_size_for(static_cast<uint32_t>(tris.size()));
            gridP.build(bP, std::move(plas), gP);
            gridS.build(bS, std::move(sphs), gS);
            gridT.build(bT, std::move(tris), gT);
        } else {
            bvhP.build(std::move(plas), leaf);
            bvhS.build(std::move(sphs), leaf);
            bvhT.build(std::move(tris), leaf);
        }
    }

    // Scene::traceLights, Scene.cpp:31-37 with AreaLight::intersect
    // (AreaLight.cpp:37-43) and PointLight::intersect (always false).
    bool trace_lights(Isect<R>* is, const Ray<R>& ray, TraceCounters* tc) const {
        bool hit = false;
        for (const Light<R>& l : lights) {
            if (l.kind != MRT_LIGHT_AREA) continue;
            if (tc) ++tc->prims;
            const bool h = l.tri.intersect(is, ray);
            if (h) is->material = &l.radiance;
            hit |= h;
        }
        return hit;
    }

    // The intersection half of Shader::rayTrace, Shader.cpp:124-156.
    bool closest(Isect<R>* is, const Ray<R>& ray, TraceCounters* tc) const {
        bool hit = false;
        if (accel == MRT_NAIVE) {
            hit |= naiveP.trace(is, ray, tc);
            hit |= naiveS.trace(is, ray, tc);
            hit |= naiveT.trace(is, ray, tc);
        } else if (accel == MRT_REGULAR_GRID) {
            hit |= gridP.trace(is, ray, tc);
            hit |= gridS.trace(is, ray, tc);
            hit |= gridT.trace(is, ray, tc);
        } else {
            hit |= bvhP.trace(is, ray, tc);
            hit |= bvhS.trace(is, ray, tc);
            hit |= bvhT.trace(is, ray, tc);
        }
        hit |= trace_lights(is, ray, tc);
        if (hit && is->materialId >= 0) is->material = &mats[static_cast<size_t>(is->materialId)];
        return hit;
    }
    // Shader::shadowTrace, Shader.cpp:162-187.
    bool shadow(const Ray<R>& ray, R dist, TraceCounters* tc) const {
        bool hit = false;
        if (accel == MRT_NAIVE) {
            hit |= naiveP.shadow(ray, dist, tc);
            hit |= naiveS.shadow(ray, dist, tc);
            hit |= naiveT.shadow(ray, dist, tc);
        } else if (accel == MRT_REGULAR_GRID) {
            hit |= gridP.shadow(ray, dist, tc);
            hit |= gridS.shadow(ray, dist, tc);
            hit |= gridT.shadow(ray, dist, tc);
        } else {
            hit |= bvhP.shadow(ray, dist, tc);
            hit |= bvhS.shadow(ray, dist, tc);
            hit |= bvhT.shadow(ray, dist, tc);
        }
        return hit;
    }
};

// ------------------------------------------------------------------ Shader --
template <class R>
struct Shader {
    const Scene<R>* scene = nullptr;
    const mrt_render_desc* desc = nullptr;
    Counters<R>* counters = nullptr;
    OrderedCounts* ordered = nullptr;  // optional: ordered-traversal accounting of every counted ray

    void account_closest(const Ray<R>& ray) const {
        if (!ordered || scene->accel != MRT_BVH) return;
        uint64_t p = 0, t = 0, lp = 0, lt = 0;
        Isect<R> is;
        is.length = static_cast<R>(kRayLengthMax);
        ordered_walk(scene->bvhT, ray, &is, false, static_cast<R>(0), &p, &t);
        ordered_walk(scene->lightBvh, ray, &is, false, static_cast<R>(0), &lp, &lt);
        ordered->rays_closest.fetch_add(1, std::memory_order_relaxed);
        ordered->pairs_closest.fetch_add(p, std::memory_order_relaxed);
        ordered->tests_closest.fetch_add(t, std::memory_order_relaxed);
        ordered->pairs_lights.fetch_add(lp, std::memory_order_relaxed);
        ordered->tests_lights.fetch_add(lt, std::memory_order_relaxed);
    }
    void account_shadow(const Ray<R>& ray, R dist) const {
        if (!ordered || scene->accel != MRT_BVH) return;
        uint64_t p = 0, t = 0;
        ordered_walk(scene->bvhT, ray, static_cast<Isect<R>*>(nullptr), true, dist, &p, &t);
        ordered->rays_shadow.fetch_add(1, std::memory_order_relaxed);
        ordered->pairs_shadow.fetch_add(p, std::memory_order_relaxed);
        ordered->tests_shadow.fetch_add(t, std::memory_order_relaxed);
    }

    float draw(const mrt_sampler& s, const DrawCtx& c, uint32_t slot, uint32_t ls, int32_t depth) const {
        return orc::draw(s, c.pixel, c.sample, make_dim(slot, ls, static_cast<uint32_t>(depth), c.branch),
                         desc->samples_pixel);
    }

    // Shader::getLight, Shader.cpp:225-236: floor(r * sizeLights * 0.99999f).
    uint32_t choose_light(const DrawCtx& c, uint32_t ls, int32_t depth) const {
        const float r = draw(desc->sampler_light, c, SLOT_LIGHT, ls, depth);
        const uint32_t n = static_cast<uint32_t>(scene->lights.size());
        const float x = (r * static_cast<float>(n)) * 0.99999f;
        return static_cast<uint32_t>(std::floor(x));
    }
    // Light::getPosition: AreaLight.cpp:22-31, PointLight.cpp:17-19.
    V3<R> light_position(const Light<R>& l, const DrawCtx& c, uint32_t ls, int32_t depth) const {
        if (l.kind != MRT_LIGHT_AREA) return l.position;
        float Rr = draw(l.sampler, c, SLOT_LIGHT_R, ls, depth);
        float Ss = draw(l.sampler, c, SLOT_LIGHT_S, ls, depth);
        if (Rr + Ss >= 1.0f) {
            Rr = 1.0f - Rr;
            Ss = 1.0f - Ss;
        }
        return (l.tri.A + static_cast<R>(Rr) * l.tri.AB) + static_cast<R>(Ss) * l.tri.AC;
    }
    // Shader::getCosineSampleHemisphere, Shader.cpp:193-223.  cos/sin are
    // evaluated in double and rounded, the GPU does the same.
    V3<R> cosine_hemisphere(const V3<R>& n, const DrawCtx& c, int32_t depth) const {
        const float r1 = draw(desc->sampler_hemi, c, SLOT_HEMI_1, 0, depth);
        const float r2 = draw(desc->sampler_hemi, c, SLOT_HEMI_2, 0, depth);
        const R phi = static_cast<R>(6.28318530717958647692f) * static_cast<R>(r1);
        const R rr2 = static_cast<R>(r2);
        const R cosTheta = std::sqrt(rr2);
        V3<R> u = std::abs(n.x) > static_cast<R>(0.1f) ? V3<R>{0, 1, 0} : V3<R>{1, 0, 0};
        u = normalize(cross(u, n));
        const V3<R> v = cross(n, u);
        const R cphi = static_cast<R>(std::cos(static_cast<double>(phi)));
        const R sphi = static_cast<R>(std::sin(static_cast<double>(phi)));
        V3<R> dir = (u * (cphi * cosTheta) + v * (sphi * cosTheta)) + n * std::sqrt(static_cast<R>(1) - rr2);
        return normalize(dir);
    }

    // Direct light loop shared by Whitted.cpp:53-82 and PathTracer.cpp:71-103.
    V3<R> direct(const Isect<R>& is, int32_t depth, const DrawCtx& c) const {
        V3<R> L{};
        const uint32_t spl = desc->samples_light;
        for (uint32_t i = 0; i < spl; ++i) {
            const Light<R>& light = scene->lights[choose_light(c, i, depth)];
            const V3<R> lp = light_position(light, c, i, depth);
            V3<R> toLight = lp - is.point;
            const R dist = length(toLight);
            toLight = normalize(toLight);
            const R cosNL = dot(is.normal, toLight);
            if (cosNL > 0) {
                Ray<R> sr;
                sr.o = is.point;
                sr.d = toLight;
                sr.depth = depth + 1;
                sr.exclude = is.prim;
                if (counters) counters->shadow.fetch_add(1, std::memory_order_relaxed);
                account_shadow(sr, dist);
                if (!scene->shadow(sr, dist, nullptr)) L = L + light.radiance.Le * cosNL;
            }
        }
        return L;
    }

    // Shader::rayTrace, Shader.cpp:124-160, dispatching to shade().
    bool ray_trace(V3<R>* rgb, const Ray<R>& ray, const DrawCtx& c, int kind) const {
        Isect<R> is;
        const bool hit = scene->closest(&is, ray, nullptr);
        if (ray.depth <= desc->ray_depth_max) account_closest(ray);
        if (counters && ray.depth <= desc->ray_depth_max) {
            if (kind == 0) counters->camera.fetch_add(1, std::memory_order_relaxed);
            else if (kind == 1) counters->extension.fetch_add(1, std::memory_order_relaxed);
            else counters->specular.fetch_add(1, std::memory_order_relaxed);
        }
        if (!hit) return false;
        // Shader::rayTrace resolves the material, Shader.cpp:151-155, then
        // calls the shader's shade().
        switch (desc->shader) {
            case MRT_SHADER_WHITTED: return whitted(rgb, is, ray, c);
            case MRT_SHADER_NOSHADOWS: return no_shadows(rgb, is, ray, c);
            case MRT_SHADER_DEPTHMAP: return depth_map(rgb, is, ray);
            case MRT_SHADER_DIFFUSE: return diffuse_material(rgb, is);
            default: return path_tracer(rgb, is, ray, c);
        }
    }

    // NoShadows::shade, app/Components/Shaders/NoShadows.cpp:19-64: the light
    // loop of Whitted without the shadow ray, then the 0.1 * kD ambient.
    bool no_shadows(V3<R>* rgb, const Isect<R>& is, const Ray<R>& ray, const DrawCtx& c) const {
        V3<R> kD{};
        if (is.material != nullptr) {
            if (any_gt0(is.material->Le)) {
                *rgb = is.material->Le;
                return true;
            }
            kD = is.material->Kd;
        } else {
            kD = is.color;
        }
        if (any_gt0(kD)) {
            if (!scene->lights.empty()) {
                const uint32_t spl = desc->samples_light;
                for (uint32_t j = 0; j < spl; ++j) {
                    const Light<R>& light = scene->lights[choose_light(c, j, ray.depth)];
                    const V3<R> lp = light_position(light, c, j, ray.depth);
                    const V3<R> toLight = normalize(lp - is.point);
                    const R cosNL = dot(is.normal, toLight);
                    if (cosNL > 0) *rgb = *rgb + light.radiance.Le * cosNL;
                }
                *rgb = *rgb * kD;
                *rgb = *rgb / static_cast<R>(spl);
            }
        }
        *rgb = *rgb + kD * static_cast<R>(0.1f);
        return false;
    }

    // DepthMap::shade, app/Components/Shaders/DepthMap.cpp:19-26.
    bool depth_map(V3<R>* rgb, const Isect<R>& is, const Ray<R>& ray) const {
        const V3<R> maxPoint{static_cast<R>(desc->depth_max_point[0]), static_cast<R>(desc->depth_max_point[1]),
                             static_cast<R>(desc->depth_max_point[2])};
        const R maxDist = length(maxPoint - ray.o) * static_cast<R>(2.0f);
        const R depth = std::max((maxDist - is.length) / maxDist, static_cast<R>(0));
        *rgb = V3<R>{depth, depth, depth};
        return false;
    }

    // DiffuseMaterial::shade, app/Components/Shaders/DiffuseMaterial.cpp:16-38.
    bool diffuse_material(V3<R>* rgb, const Isect<R>& is) const {
        if (is.material != nullptr) {
            if (any_gt0(is.material->Kd)) *rgb = is.material->Kd;
            else if (any_gt0(is.material->Ks)) *rgb = is.material->Ks;
            else if (any_gt0(is.material->Kt)) *rgb = is.material->Kt;
            else if (any_gt0(is.material->Le)) *rgb = is.material->Le;
        } else {
            *rgb = is.color;
        }
        return false;
    }

    // Whitted::shade, app/Components/Shaders/Whitted.cpp:17-107.
    bool whitted(V3<R>* rgb, const Isect<R>& is, const Ray<R>& ray, const DrawCtx& c) const {
        const int32_t depth = ray.depth;
        if (depth > desc->ray_depth_max) return false;
        V3<R> kD{}, kS{}, kT{};
        R ior = 1;
        if (is.material != nullptr) {
            if (any_gt0(is.material->Le)) {
                *rgb = is.material->Le;
                return true;
            }
            kD = is.material->Kd;
            kS = is.material->Ks;
            kT = is.material->Kt;
            ior = is.material->ior;
        } else {
            kD = is.color;
        }
        if (any_gt0(kD)) {
            if (!scene->lights.empty()) {
                *rgb = *rgb + direct(is, depth, c);
                *rgb = *rgb * kD;
                *rgb = *rgb / static_cast<R>(desc->samples_light);
            }
        }
        if (any_gt0(kS)) {
            Ray<R> sr;
            sr.o = is.point;
            sr.d = reflect(ray.d, is.normal);
            sr.depth = depth + 1;
            sr.exclude = is.prim;
            DrawCtx cc = c;
            cc.branch = c.branch * 3u + 1u;
            V3<R> li{};
            ray_trace(&li, sr, cc, 2);
            *rgb = *rgb + kS * li;
        }
        if (any_gt0(kT)) {
            Ray<R> tr;
            tr.o = is.point;
            tr.d = refract(ray.d, is.normal, static_cast<R>(1) / ior);
            tr.depth = depth + 1;
            tr.exclude = is.prim;
            DrawCtx cc = c;
            cc.branch = c.branch * 3u + 2u;
            V3<R> li{};
            ray_trace(&li, tr, cc, 2);
            *rgb = *rgb + kT * li;
        }
        *rgb = *rgb + kD * static_cast<R>(0.1f);
        return false;
    }

    // PathTracer::shade, app/Components/Shaders/PathTracer.cpp:25-164, with
    // RayDepthMin/RayDepthMax taken from the render descriptor.
    bool path_tracer(V3<R>* rgb, const Isect<R>& is, const Ray<R>& ray, const DrawCtx& c) const {
        const int32_t depth = ray.depth;
        if (depth > desc->ray_depth_max) return false;
        V3<R> kD{}, kS{}, kT{};
        R ior = 1;
        if (is.material != nullptr) {
            if (any_gt0(is.material->Le)) {
                *rgb = is.material->Le;
                return true;
            }
            kD = is.material->Kd;
            kS = is.material->Ks;
            kT = is.material->Kt;
            ior = is.material->ior;
        } else {
            kD = is.color;
        }
        V3<R> Ld{}, LiD{}, LiS{}, LiT{};
        bool hitLight = false;
        if (any_gt0(kD)) {
            if (!scene->lights.empty()) {
                Ld = direct(is, depth, c);
                Ld = Ld * kD;
                Ld = Ld / static_cast<R>(desc->samples_light);
            }
            bool cont = depth <= desc->ray_depth_min;
            if (!cont) cont = draw(desc->sampler_rr, c, SLOT_RR, 0, depth) > 0.5f;
            if (cont) {
                Ray<R> nr;
                nr.o = is.point;
                nr.d = cosine_hemisphere(is.normal, c, depth);
                nr.depth = depth + 1;
                nr.exclude = is.prim;
                V3<R> li{};
                hitLight = ray_trace(&li, nr, c, 1);
                LiD = LiD + kD * li;
                if (depth > desc->ray_depth_min) LiD = LiD / (static_cast<R>(0.5f) * static_cast<R>(0.5f));
                if (any_gt0(Ld) && hitLight) LiD = V3<R>{};
            }
        }
        if (any_gt0(kS)) {
            Ray<R> sr;
            sr.o = is.point;
            sr.d = reflect(ray.d, is.normal);
            sr.depth = depth + 1;
            sr.exclude = is.prim;
            DrawCtx cc = c;
            cc.branch = c.branch * 3u + 1u;
            V3<R> li{};
            ray_trace(&li, sr, cc, 2);
            LiS = LiS + kS * li;
        }
        if (any_gt0(kT)) {
            Ray<R> tr;
            tr.o = is.point;
            tr.d = refract(ray.d, is.normal, static_cast<R>(1) / ior);
            tr.depth = depth + 1;
            tr.exclude = is.prim;
            DrawCtx cc = c;
            cc.branch = c.branch * 3u + 2u;
            V3<R> li{};
            ray_trace(&li, tr, cc, 2);
            LiT = LiT + kT * li;
        }
        *rgb = *rgb + Ld;
        *rgb = *rgb + LiD;
        *rgb = *rgb + LiS;
        *rgb = *rgb + LiT;
        return hitLight;
    }
};

// ------------------------------------------------------------------ Camera --
// Perspective::fastArcTan, app/Components/Cameras/Perspective.cpp:39-44.
inline float fast_arctan(float value) {
    const float absValue = std::abs(value);
    const float a = 0.785398163397448309615660845819875721f * value;
    const float b = value * (absValue - 1.0f);
    const float c = 0.2447f + 0.0663f * absValue;
    return a - b * c;
}
template <class R>
Ray<R> generate_ray(const mrt_camera& cam, float u, float v, float du, float dv) {
    const V3<R> pos = from3<R>(cam.position), dir = from3<R>(cam.direction);
    const V3<R> right = from3<R>(cam.right), up = from3<R>(cam.up);
    Ray<R> r;
    r.depth = 1;
    r.exclude = MRT_ID_NONE;
    if (cam.kind == MRT_CAMERA_PERSPECTIVE) {
        // Perspective::generateRay, Perspective.cpp:24-36.
        const float tvr = cam.fov_h * (u - 0.5f);
        const float rf = fast_arctan(tvr) + du;
        const V3<R> rr = right * static_cast<R>(rf);
        const float tvu = cam.fov_v * (0.5f - v);
        const float uf = fast_arctan(tvu) + dv;
        const V3<R> uu = up * static_cast<R>(uf);
        const V3<R> dest = ((pos + dir) + rr) + uu;
        r.o = pos;
        r.d = normalize(dest - pos);
    } else {
        // Orthographic::generateRay, Orthographic.cpp:23-31.
        const float rf = (u - 0.5f) * cam.fov_h;
        const V3<R> rr = right * static_cast<R>(rf) + right * static_cast<R>(du);
        const float uf = (0.5f - v) * cam.fov_v;
        const V3<R> uu = up * static_cast<R>(uf) + up * static_cast<R>(dv);
        r.o = (pos + rr) + uu;
        r.d = dir;
    }
    return r;
}

// incrementalAvg, app/MobileRT/Utils.cpp:48-71.
template <class R>
uint32_t incremental_avg(const V3<R>& s, uint32_t avg, uint32_t n) {
    const uint32_t lr = avg & 0xFFu, lg = (avg >> 8u) & 0xFFu, lb = (avg >> 16u) & 0xFFu;
    const uint32_t sr = static_cast<uint32_t>(static_cast<int64_t>(s.x * static_cast<R>(255)));
    const uint32_t sg = static_cast<uint32_t>(static_cast<int64_t>(s.y * static_cast<R>(255)));
    const uint32_t sb = static_cast<uint32_t>(static_cast<int64_t>(s.z * static_cast<R>(255)));
    const uint32_t cr = ((n - 1) * lr + sr) / n;
    const uint32_t cg = ((n - 1) * lg + sg) / n;
    const uint32_t cb = ((n - 1) * lb + sb) / n;
    return 0xFF000000u | (std::min(cb, 255u) << 16u) | (std::min(cg, 255u) << 8u) | std::min(cr, 255u);
}

// ---------------------------------------------------------------- Renderer --
struct RenderArgs {
    uint32_t* bitmap;
    uint32_t stride_bytes;
    double* accum;            // optional, 3 per pixel (width stride)
    uint32_t nthreads;
    const uint8_t* tile_mask; // optional: only tiles (claim rank) with mask != 0
    uint32_t pixel_mod = 0;   // optional: only pixels whose id is a multiple of it
    OrderedCounts* ordered = nullptr;
};

// Renderer::renderFrame + renderScene, app/MobileRT/Renderer.cpp:31-116.
// Tiles are claimed from an atomic counter exactly like Camera::getBlock
// (Camera.cpp:28-37); a barrier separates sample passes so results do not
// depend on the interleaving (the reference has none).
template <class R>
int render(const Scene<R>& scene, const mrt_render_desc& d, const RenderArgs& a, uint64_t cnt[4]) {
    const uint32_t w = d.width, h = d.height;
    if (w < 16 || h < 16) return MRT_E_INVALID;
    const uint32_t bsx = w / 16u, bsy = h / 16u;
    const uint32_t domain = (w / bsx) * (h / bsy);
    const uint32_t resolution = w * h;
    const uint32_t realWidth = a.stride_bytes / 4u;
    const float invW = 1.0f / static_cast<float>(w);
    const float invH = 1.0f / static_cast<float>(h);
    const float pixW = 0.5f / static_cast<float>(w);
    const float pixH = 0.5f / static_cast<float>(h);
    Counters<R> counters;
    Shader<R> sh;
    sh.scene = &scene;
    sh.desc = &d;
    sh.counters = &counters;
    sh.ordered = a.ordered;
    std::atomic<uint32_t> compat{0};  // Sampler::sample_, reset per frame
    const bool isCompat = d.sampler_pixel.kind == MRT_SAMPLER_HALTON_COMPAT ||
                          d.sampler_pixel.kind == MRT_SAMPLER_STRATIFIED_COMPAT;
    auto pixelDraw = [&](uint32_t pixel, uint32_t sample, uint32_t slot) -> float {
        if (!isCompat) {
            return orc::draw(d.sampler_pixel, pixel, sample, make_dim(slot, 0, 1, 0), d.samples_pixel);
        }
        // HaltonSeq/Stratified::getSample(0), HaltonSeq.cpp:14-22, Stratified.cpp:14-23.
        const uint32_t D = d.sampler_pixel.domain_size;
        const uint32_t cur = compat.fetch_add(1, std::memory_order_relaxed);
        if (cur >= D) {
            compat.fetch_sub(1, std::memory_order_relaxed);
            return 1.0f;
        }
        if (d.sampler_pixel.kind == MRT_SAMPLER_HALTON_COMPAT) return halton_sequence(cur, 2u);
        return static_cast<float>(cur) / static_cast<float>(D);
    };
    for (uint32_t sample = d.first_sample; sample < d.samples_pixel; ++sample) {
        std::atomic<uint32_t> block{0};
        auto worker = [&]() {
            while (true) {
                const uint32_t t = block.fetch_add(1, std::memory_order_relaxed);
                if (t >= kNumberOfBlocks) break;
                if (a.tile_mask && !a.tile_mask[t]) continue;
                const float blockV = d.tile_values[t];
                const uint32_t roundBlock = static_cast<uint32_t>(std::roundf(blockV * static_cast<float>(domain)));
                const uint32_t pixel = roundBlock * bsx % resolution;
                const uint32_t startY = ((pixel / w) * bsy) % h;
                const uint32_t endY = startY + bsy;
                // tile_skip (no counterpart in the reference): the tile is
                // not rendered; a compat counter still advances by the
                // tile's draws, so every other pixel draws what it draws in
                // the whole frame.
                const uint32_t tileId = (startY / bsy) * 16u + (pixel % w) / bsx;
                if ((d.tile_skip[tileId >> 5] >> (tileId & 31u)) & 1u) {
                    if (isCompat)
                        for (uint32_t k = 0; k < 2u * bsx * bsy; ++k) (void)pixelDraw(0, sample, SLOT_JITTER_U);
                    continue;
                }
                for (uint32_t y = startY; y < endY; ++y) {
                    const float v = static_cast<float>(y) * invH;
                    const uint32_t yWidth = y * realWidth;
                    const uint32_t startX = (pixel + yWidth) % w;
                    const uint32_t endX = startX + bsx;
                    for (uint32_t x = startX; x < endX; ++x) {
                        const float u = static_cast<float>(x) * invW;
                        const uint32_t pid = y * w + x;
                        if (a.pixel_mod && pid % a.pixel_mod != 0u) continue;
                        const float r1 = pixelDraw(pid, sample, SLOT_JITTER_U);
                        const float r2 = pixelDraw(pid, sample, SLOT_JITTER_V);
                        const float du = ((r1 - 0.5f) * 2.0f) * pixW;
                        const float dv = ((r2 - 0.5f) * 2.0f) * pixH;
                        const Ray<R> ray = generate_ray<R>(d.camera, u, v, du, dv);
                        V3<R> rgb{};
                        DrawCtx c;
                        c.pixel = pid;
                        c.sample = sample;
                        c.branch = 0;
                        sh.ray_trace(&rgb, ray, c, 0);
                        const uint32_t bi = yWidth + x;
                        a.bitmap[bi] = incremental_avg(rgb, a.bitmap[bi], sample + 1);
                        if (a.accum) {
                            a.accum[3 * static_cast<size_t>(pid) + 0] += static_cast<double>(rgb.x);
                            a.accum[3 * static_cast<size_t>(pid) + 1] += static_cast<double>(rgb.y);
                            a.accum[3 * static_cast<size_t>(pid) + 2] += static_cast<double>(rgb.z);
                        }
                    }
                }
            }
        };
        const uint32_t nt = std::max(1u, a.nthreads);
        std::vector<std::thread> ts;
        for (uint32_t i = 1; i < nt; ++i) ts.emplace_back(worker);
        worker();
        for (auto& t : ts) t.join();
    }
    cnt[0] = counters.camera.load();
    cnt[1] = counters.extension.load();
    cnt[2] = counters.shadow.load();
    cnt[3] = counters.specular.load();
    return MRT_OK;
}

// Holder of the float and double scenes built from the same stored forms.
struct Handle {
    std::unique_ptr<Scene<float>> f;
    std::unique_ptr<Scene<double>> d;
};

template <class R>
std::unique_ptr<Scene<R>> make_scene(const mrt_scene_desc* s) {
    auto sc = std::make_unique<Scene<R>>();
    for (uint32_t i = 0; i < s->n_triangles; ++i)
        sc->tris.push_back(Triangle<R>::from(s->triangles[i], MRT_MAKE_ID(MRT_CAT_TRIANGLE, i)));
    for (uint32_t i = 0; i < s->n_spheres; ++i)
        sc->sphs.push_back(Sphere<R>::from(s->spheres[i], MRT_MAKE_ID(MRT_CAT_SPHERE, i)));
    for (uint32_t i = 0; i < s->n_planes; ++i)
        sc->plas.push_back(Plane<R>::from(s->planes[i], MRT_MAKE_ID(MRT_CAT_PLANE, i)));
    for (uint32_t i = 0; i < s->n_materials; ++i) sc->mats.push_back(mat_from<R>(s->materials[i]));
    for (uint32_t i = 0; i < s->n_lights; ++i) {
        Light<R> l;
        l.kind = s->lights[i].kind;
        l.radiance = mat_from<R>(s->lights[i].radiance);
        l.tri = Triangle<R>::from(s->lights[i].triangle, MRT_MAKE_ID(MRT_CAT_LIGHT, i));
        l.position = from3<R>(s->lights[i].position);
        l.sampler = s->lights[i].sampler;
        sc->lights.push_back(l);
    }
    return sc;
}

template <class R>
void fill_hit(const Isect<R>& is, bool hit, mrt_hit* out) {
    if (!hit) {
        out->t = 0.0f;
        out->id = MRT_ID_NONE;
        out->u = 0.0f;
        out->v = 0.0f;
        return;
    }
    out->t = static_cast<float>(is.length);
    out->id = is.hitId;
    out->u = static_cast<float>(is.u);
    out->v = static_cast<float>(is.v);
}

}  // namespace orc

using namespace orc;

extern "C" {

// ------------------------------------------------------------ scene API --
void* oracle_scene_create(const mrt_scene_desc* desc) {
    auto* h = new Handle;
    h->f = make_scene<float>(desc);
    h->d = make_scene<double>(desc);
    return h;
}
void oracle_scene_destroy(void* h) { delete static_cast<Handle*>(h); }

int oracle_scene_build(void* hv, int accel, const mrt_accel_params* p) {
    auto* h = static_cast<Handle*>(hv);
    const uint32_t leaf = p && p->leaf_size ? p->leaf_size : 2u;
    const float zero[3] = {0, 0, 0};
    const float* mn = p ? p->camera_min : zero;
    const float* mx = p ? p->camera_max : zero;
    h->f->build(accel, from3<float>(mn), from3<float>(mx), leaf);
    h->d->build(accel, from3<double>(mn), from3<double>(mx), leaf);
    return 0;
}

// Counts interior nodes / returns the BVH node array of one category (float).
uint32_t oracle_bvh_nodes(void* hv, int category, BVHNodeOut* out, uint32_t cap) {
    auto* h = static_cast<Handle*>(hv);
    auto dump = [&](const auto& bvh) -> uint32_t {
        const uint32_t n = static_cast<uint32_t>(bvh.boxes.size());
        if (out) {
            for (uint32_t i = 0; i < n && i < cap; ++i) {
                const auto& nd = bvh.boxes[i];
                out[i].mn[0] = nd.box.mn.x; out[i].mn[1] = nd.box.mn.y; out[i].mn[2] = nd.box.mn.z;
                out[i].mx[0] = nd.box.mx.x; out[i].mx[1] = nd.box.mx.y; out[i].mx[2] = nd.box.mx.z;
                out[i].indexOffset = nd.indexOffset;
                out[i].numberPrimitives = nd.numberPrimitives;
            }
        }
        return n;
    };
    if (category == MRT_CAT_PLANE) return dump(h->f->bvhP);
    if (category == MRT_CAT_SPHERE) return dump(h->f->bvhS);
    return dump(h->f->bvhT);
}

// Grid description of one category: size, bounds, and the CSR cell lists.
int oracle_grid_info(void* hv, int category, int32_t* grid_size, float bounds[6], uint32_t* n_refs) {
    auto* h = static_cast<Handle*>(hv);
    auto info = [&](const auto& g) {
        *grid_size = g.gridSize;
        bounds[0] = g.ext.mn.x; bounds[1] = g.ext.mn.y; bounds[2] = g.ext.mn.z;
        bounds[3] = g.ext.mx.x; bounds[4] = g.ext.mx.y; bounds[5] = g.ext.mx.z;
        uint64_t n = 0;
        for (const auto& c : g.cells) n += c.size();
        *n_refs = static_cast<uint32_t>(n);
        return 0;
    };
    if (category == MRT_CAT_PLANE) return info(h->f->gridP);
    if (category == MRT_CAT_SPHERE) return info(h->f->gridS);
    return info(h->f->gridT);
}
int oracle_grid_cells(void* hv, int category, uint32_t* offsets, uint32_t* refs) {
    auto* h = static_cast<Handle*>(hv);
    auto csr = [&](const auto& g) {
        uint32_t k = 0;
        for (size_t i = 0; i < g.cells.size(); ++i) {
            offsets[i] = k;
            for (uint32_t r : g.cells[i]) refs[k++] = r;
        }
        offsets[g.cells.size()] = k;
        return 0;
    };
    if (category == MRT_CAT_PLANE) return csr(h->f->gridP);
    if (category == MRT_CAT_SPHERE) return csr(h->f->gridS);
    return csr(h->f->gridT);
}

static Ray<float> ray_f(const mrt_ray& r) {
    Ray<float> x;
    x.o = from3<float>(r.org);
    x.d = from3<float>(r.dir);
    x.depth = 1;
    x.exclude = r.exclude;
    return x;
}

// Shader::rayTrace intersection step for a batch (multi-threaded).
int oracle_trace_closest(void* hv, const mrt_ray* rays, uint32_t n, mrt_hit* hits, uint32_t nthreads,
                         uint64_t* pairs_out, uint64_t* prims_out) {
    auto* h = static_cast<Handle*>(hv);
    std::atomic<uint64_t> pairs{0}, prims{0};
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        TraceCounters tc;
        while (true) {
            const uint32_t i0 = next.fetch_add(256);
            if (i0 >= n) break;
            const uint32_t i1 = std::min(n, i0 + 256);
            for (uint32_t i = i0; i < i1; ++i) {
                Isect<float> is;
                is.length = rays[i].tmax;
                const bool hit = h->f->closest(&is, ray_f(rays[i]), &tc);
                fill_hit(is, hit, &hits[i]);
            }
        }
        pairs += tc.pairs;
        prims += tc.prims;
    };
    const uint32_t nt = std::max(1u, nthreads);
    std::vector<std::thread> ts;
    for (uint32_t i = 1; i < nt; ++i) ts.emplace_back(work);
    work();
    for (auto& t : ts) t.join();
    if (pairs_out) *pairs_out = pairs.load();
    if (prims_out) *prims_out = prims.load();
    return 0;
}

int oracle_trace_any(void* hv, const mrt_ray* rays, uint32_t n, uint8_t* occ, uint32_t nthreads) {
    auto* h = static_cast<Handle*>(hv);
    std::atomic<uint32_t> next{0};
    auto work = [&]() {
        while (true) {
            const uint32_t i0 = next.fetch_add(256);
            if (i0 >= n) break;
            const uint32_t i1 = std::min(n, i0 + 256);
            for (uint32_t i = i0; i < i1; ++i) occ[i] = h->f->shadow(ray_f(rays[i]), rays[i].tmax, nullptr) ? 1 : 0;
        }
    };
    const uint32_t nt = std::max(1u, nthreads);
    std::vector<std::thread> ts;
    for (uint32_t i = 1; i < nt; ++i) ts.emplace_back(work);
    work();
    for (auto& t : ts) t.join();
    return 0;
}

// Shader::rayTrace + shade for a batch of rays (ray i = pixel i).
int oracle_shade_rays(void* hv, const mrt_render_desc* d, const mrt_ray* rays, const int32_t* depths,
                      uint32_t n, uint32_t sample, float* rgb, uint8_t* hit_light, int use_double) {
    auto* h = static_cast<Handle*>(hv);
    for (uint32_t i = 0; i < n; ++i) {
        DrawCtx c;
        c.pixel = i;
        c.sample = sample;
        if (use_double) {
            Shader<double> sh;
            sh.scene = h->d.get();
            sh.desc = d;
            Ray<double> r;
            r.o = from3<double>(rays[i].org);
            r.d = from3<double>(rays[i].dir);
            r.depth = depths ? depths[i] : 1;
            r.exclude = rays[i].exclude;
            V3<double> out{};
            const bool hl = sh.ray_trace(&out, r, c, 0);
            rgb[3 * i] = static_cast<float>(out.x);
            rgb[3 * i + 1] = static_cast<float>(out.y);
            rgb[3 * i + 2] = static_cast<float>(out.z);
            if (hit_light) hit_light[i] = hl ? 1 : 0;
        } else {
            Shader<float> sh;
            sh.scene = h->f.get();
            sh.desc = d;
            Ray<float> r = ray_f(rays[i]);
            r.depth = depths ? depths[i] : 1;
            V3<float> out{};
            const bool hl = sh.ray_trace(&out, r, c, 0);
            rgb[3 * i] = out.x;
            rgb[3 * i + 1] = out.y;
            rgb[3 * i + 2] = out.z;
            if (hit_light) hit_light[i] = hl ? 1 : 0;
        }
    }
    return 0;
}

// Renderer::renderFrame.  accum (double, 3 per pixel) is optional.
int oracle_render(void* hv, const mrt_render_desc* d, uint32_t* bitmap, uint32_t stride_bytes, double* accum,
                  uint32_t nthreads, const uint8_t* tile_mask, int use_double, uint64_t counters[4],
                  double* seconds) {
    auto* h = static_cast<Handle*>(hv);
    RenderArgs a{bitmap, stride_bytes, accum, nthreads, tile_mask};
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    if (use_double) rc = render<double>(*h->d, *d, a, counters);
    else rc = render<float>(*h->f, *d, a, counters);
    const auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    return rc;
}

// Traffic of an ordered traversal of the scene's trees (measurement): renders
// the pixels whose id is a multiple of pixel_mod and, for every counted ray,
// walks the same reference BVH near-first with culling (closest hit: the
// triangle tree, then a reference-built tree over the area lights; any hit:
// the triangle tree).  out: closest-hit rays, their pair visits, triangle
// tests; shadow rays, pair visits, tests; light-tree pair visits, tests.
int oracle_ordered_traffic(void* hv, const mrt_render_desc* d, uint32_t pixel_mod, uint32_t nthreads, double out[8]) {
    auto* h = static_cast<Handle*>(hv);
    Scene<float>& sc = *h->f;
    if (sc.accel != MRT_BVH) return MRT_E_INVALID;
    if (sc.lightBvh.boxes.empty()) {
        std::vector<Triangle<float>> lt;
        for (const auto& l : sc.lights)
            if (l.kind == MRT_LIGHT_AREA) lt.push_back(l.tri);
        sc.lightBvh.build(std::move(lt), 2);
    }
    std::vector<uint32_t> bitmap(static_cast<size_t>(d->width) * d->height, 0u);
    OrderedCounts oc;
    RenderArgs a{bitmap.data(), d->width * 4u, nullptr, nthreads, nullptr, pixel_mod, &oc};
    uint64_t cnt[4] = {};
    const int rc = render<float>(sc, *d, a, cnt);
    out[0] = static_cast<double>(oc.rays_closest.load());
    out[1] = static_cast<double>(oc.pairs_closest.load());
    out[2] = static_cast<double>(oc.tests_closest.load());
    out[3] = static_cast<double>(oc.rays_shadow.load());
    out[4] = static_cast<double>(oc.pairs_shadow.load());
    out[5] = static_cast<double>(oc.tests_shadow.load());
    out[6] = static_cast<double>(oc.pairs_lights.load());
    out[7] = static_cast<double>(oc.tests_lights.load());
    return rc;
}

int oracle_camera_rays(const mrt_camera* cam, const float* uvdd, uint32_t n, mrt_ray* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const Ray<float> r = generate_ray<float>(*cam, uvdd[4 * i], uvdd[4 * i + 1], uvdd[4 * i + 2], uvdd[4 * i + 3]);
        out[i].org[0] = r.o.x; out[i].org[1] = r.o.y; out[i].org[2] = r.o.z;
        out[i].dir[0] = r.d.x; out[i].dir[1] = r.d.y; out[i].dir[2] = r.d.z;
        out[i].tmax = kRayLengthMax;
        out[i].exclude = MRT_ID_NONE;
    }
    return 0;
}

// ---------------------------------------------------- construction helpers --
// Triangle constructors, Triangle.cpp:12-55.
int oracle_triangle_make(const float* a, const float* b, const float* c, int32_t mat, const float* normals9,
                         const float* colors9, mrt_triangle* out) {
    const V3<float> A = from3<float>(a), B = from3<float>(b), C = from3<float>(c);
    const V3<float> AC = C - A, AB = B - A;
    std::memset(out, 0, sizeof(*out));
    auto put = [](float* dst, const V3<float>& v) { dst[0] = v.x; dst[1] = v.y; dst[2] = v.z; };
    put(out->ac, AC);
    put(out->ab, AB);
    put(out->a, A);
    if (normals9) {
        std::memcpy(out->na, normals9, 3 * sizeof(float));
        std::memcpy(out->nb, normals9 + 3, 3 * sizeof(float));
        std::memcpy(out->nc, normals9 + 6, 3 * sizeof(float));
    } else {
        const V3<float> n = normalize(cross(AC, AB));
        put(out->na, n);
        put(out->nb, n);
        put(out->nc, n);
    }
    if (colors9) {
        std::memcpy(out->ca, colors9, 3 * sizeof(float));
        std::memcpy(out->cb, colors9 + 3, 3 * sizeof(float));
        std::memcpy(out->cc, colors9 + 6, 3 * sizeof(float));
    }
    out->material_id = mat;
    return 0;
}
// Camera ctor (Camera.cpp:47-54) + Perspective/Orthographic fields.
int oracle_camera_make(int kind, const float* pos, const float* look, const float* upv, float a, float b,
                       mrt_camera* out) {
    const V3<float> P = from3<float>(pos), L = from3<float>(look), U = from3<float>(upv);
    const V3<float> dir = normalize(L - P);
    const V3<float> right = cross(U, dir);
    const V3<float> up = cross(dir, right);
    out->kind = kind;
    out->position[0] = P.x; out->position[1] = P.y; out->position[2] = P.z;
    out->direction[0] = dir.x; out->direction[1] = dir.y; out->direction[2] = dir.z;
    out->right[0] = right.x; out->right[1] = right.y; out->right[2] = right.z;
    out->up[0] = up.x; out->up[1] = up.y; out->up[2] = up.z;
    if (kind == MRT_CAMERA_PERSPECTIVE) {
        const float pi = 3.14159265358979323846264338327950288f;  // glm::pi<float>()
        out->fov_h = (a * pi) / 180.0f;                           // Camera::degToRad, Camera.cpp:72-75
        out->fov_v = (b * pi) / 180.0f;
    } else {
        out->fov_h = a / 2.0f;
        out->fov_v = b / 2.0f;
    }
    return 0;
}

// Known-answer helpers for the reference's unit tests.
int oracle_triangle_intersect(const mrt_triangle* t, const mrt_ray* r, mrt_hit* hit, float* normal_out) {
    const Triangle<float> tri = Triangle<float>::from(*t, MRT_MAKE_ID(MRT_CAT_TRIANGLE, 0));
    Isect<float> is;
    is.length = r->tmax;
    const bool h = tri.intersect(&is, ray_f(*r));
    fill_hit(is, h, hit);
    if (normal_out) {
        normal_out[0] = is.normal.x;
        normal_out[1] = is.normal.y;
        normal_out[2] = is.normal.z;
    }
    return h ? 1 : 0;
}
int oracle_triangle_intersect_any(const mrt_triangle* t, const mrt_ray* r) {
    const Triangle<float> tri = Triangle<float>::from(*t, MRT_MAKE_ID(MRT_CAT_TRIANGLE, 0));
    return tri.intersect(ray_f(*r), r->tmax) ? 1 : 0;
}
int oracle_triangle_aabb(const mrt_triangle* t, float out[6]) {
    const AABB<float> b = Triangle<float>::from(*t, 0).aabb();
    out[0] = b.mn.x; out[1] = b.mn.y; out[2] = b.mn.z;
    out[3] = b.mx.x; out[4] = b.mx.y; out[5] = b.mx.z;
    return 0;
}
int oracle_triangle_intersect_box(const mrt_triangle* t, const float* mn, const float* mx) {
    const Triangle<float> tri = Triangle<float>::from(*t, MRT_MAKE_ID(MRT_CAT_TRIANGLE, 0));
    return tri.intersect_box(AABB<float>{from3<float>(mn), from3<float>(mx)}) ? 1 : 0;
}
int oracle_sphere_intersect_box(const mrt_sphere* sp, const float* mn, const float* mx) {
    const Sphere<float> sph = Sphere<float>::from(*sp, MRT_MAKE_ID(MRT_CAT_SPHERE, 0));
    return sph.intersect_box(AABB<float>{from3<float>(mn), from3<float>(mx)}) ? 1 : 0;
}
// RegularGrid<T>::trace of the triangle grid alone for one ray: the hit and
// the cells the walk visited, in order (at most cap are written; *n_cells is
// the full count).
int oracle_grid_walk(void* hv, const mrt_ray* ray, uint32_t* cells, uint32_t cap, uint32_t* n_cells, mrt_hit* hit) {
    auto* h = static_cast<Handle*>(hv);
    if (h->f->accel != MRT_REGULAR_GRID) return MRT_E_INVALID;
    std::vector<uint32_t> visited;
    TraceCounters tc;
    tc.cells = &visited;
    Isect<float> is;
    is.length = ray->tmax;
    const bool hv2 = h->f->gridT.trace(&is, ray_f(*ray), &tc);
    fill_hit(is, hv2, hit);
    *n_cells = static_cast<uint32_t>(visited.size());
    for (uint32_t i = 0; i < visited.size() && i < cap; ++i) cells[i] = visited[i];
    return 0;
}
int oracle_plane_intersect_box(const mrt_plane* p, const float* mn, const float* mx) {
    const Plane<float> pl = Plane<float>::from(*p, MRT_MAKE_ID(MRT_CAT_PLANE, 0));
    return pl.intersect_box(AABB<float>{from3<float>(mn), from3<float>(mx)}) ? 1 : 0;
}
float oracle_plane_distance(const mrt_plane* p, const float* pt) {
    const Plane<float> pl = Plane<float>::from(*p, MRT_MAKE_ID(MRT_CAT_PLANE, 0));
    return pl.distance(from3<float>(pt));
}
int oracle_plane_make(const float* point, const float* normal, int32_t mat, mrt_plane* out) {
    const V3<float> n = normalize(from3<float>(normal));  // Plane ctor, Plane.cpp:10-15
    out->normal[0] = n.x; out->normal[1] = n.y; out->normal[2] = n.z;
    out->point[0] = point[0]; out->point[1] = point[1]; out->point[2] = point[2];
    out->material_id = mat;
    return 0;
}
int oracle_slab(const float* mn, const float* mx, const mrt_ray* r) {
    return slab(AABB<float>{from3<float>(mn), from3<float>(mx)}, from3<float>(r->org), from3<float>(r->dir)) ? 1 : 0;
}
float oracle_halton(uint32_t index, uint32_t base) { return halton_sequence(index, base); }
float oracle_fast_arctan(float v) { return fast_arctan(v); }
uint32_t oracle_incremental_avg(const float* rgb, uint32_t avg, uint32_t n) {
    return incremental_avg(V3<float>{rgb[0], rgb[1], rgb[2]}, avg, n);
}
float oracle_draw(const mrt_sampler* s, uint32_t pixel, uint32_t sample, uint32_t dim, uint32_t spp) {
    return draw(*s, pixel, sample, dim, spp);
}
int32_t oracle_grid_size_for(uint32_t n) { return grid_size_for(n); }
int oracle_struct_sizes(uint32_t* out) {
    out[0] = sizeof(mrt_triangle);
    out[1] = sizeof(mrt_sphere);
    out[2] = sizeof(mrt_plane);
    out[3] = sizeof(mrt_material);
    out[4] = sizeof(mrt_light);
    out[5] = sizeof(mrt_ray);
    out[6] = sizeof(mrt_hit);
    out[7] = sizeof(mrt_camera);
    out[8] = sizeof(mrt_render_desc);
    out[9] = sizeof(mrt_render_stats);
    out[10] = sizeof(mrt_sampler);
    out[11] = sizeof(mrt_scene_desc);
    out[12] = sizeof(mrt_server_desc);
    out[13] = sizeof(mrt_job_info);
    out[14] = sizeof(mrt_serving_stats);
    return 15;
}

}  // extern "C"
