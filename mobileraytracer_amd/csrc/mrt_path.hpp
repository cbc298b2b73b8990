// Device shading: camera rays, light sampling, the cosine hemisphere and the
// Whitted / PathTracer estimators written as an iterative path loop.
//
// The reference shaders recurse (Shader::rayTrace -> shade -> rayTrace,
// app/Components/Shaders/PathTracer.cpp:25-164, Whitted.cpp:17-107).  Here a
// pixel-sample runs as one or more chains:
//  * a chain follows the diffuse continuation of PathTracer::shade; each
//    vertex adds T * Ld to acc_all, where T is the product of kD (x4 beyond
//    RayDepthMin, PathTracer.cpp:124-126) of the vertices above it;
//  * the reference zeroes LiD at a vertex when any(Ld > 0) there and the
//    diffuse child chain ended on an emitter (PathTracer.cpp:129-131); the
//    flag travels unchanged up the chain, so the chain's value is acc_prefix
//    (the sum up to the first vertex d* with any(Ld > 0)) when the chain ends
//    on an emitter and d* exists, acc_all otherwise;
//  * reflection / transmission children (PathTracer.cpp:136-157,
//    Whitted.cpp:85-104) are queued as new chains with weight T * kS or
//    T * kT; a child queued at a vertex deeper than d* is still traced when
//    the chain's cut applies (the reference traces it too) but its value is
//    dropped, exactly the part of the tree the zeroed LiD held;
//  * a Whitted chain is one vertex: direct * kD / samplesLight + 0.1 * kD.
// Rays deeper than RayDepthMax are not traced: the reference traces them but
// shade() returns 0 and false before reading the hit (PathTracer.cpp:28-31).
#pragma once
#include "mrt_sampling.hpp"
#include "mrt_trace.hpp"

namespace mrt {

struct RenderParams {
    int32_t shader;
    uint32_t width, height;
    uint32_t spp;
    uint32_t samples_light;
    int32_t depth_min, depth_max;
    mrt_camera cam;
    mrt_sampler s_pixel, s_rr, s_light, s_hemi;
    uint32_t bsx, bsy;            // tile size (Renderer.cpp:20-21)
    uint32_t real_width;          // bitmap row stride in pixels
    float inv_w, inv_h, pix_w, pix_h;
    const uint32_t* tile_rank;    // claim rank of tile (ty * 16 + tx)
    uint32_t compat;              // pixel sampler is a *_COMPAT kind
    uint32_t pixels_per_pass;     // 256 * bsx * bsy
    float max_point[3];           // DepthMap maxPoint_
    uint32_t tile_skip[8];        // bit ty * 16 + tx: tile not rendered
};

// Ray counters of one lane, reduced per wave at the end of a launch.
struct RayCounts {
    uint32_t camera, extension, shadow, specular;
    TraceStats ts;
};

// Perspective::fastArcTan, app/Components/Cameras/Perspective.cpp:39-44.
MRT_HD float fast_arctan(float value) {
    const float absValue = fabsf(value);
    const float a = 0.785398163397448309615660845819875721f * value;
    const float b = value * (absValue - 1.0f);
    const float c = 0.2447f + 0.0663f * absValue;
    return a - b * c;
}

// Camera::generateRay: Perspective.cpp:24-36, Orthographic.cpp:23-31.
MRT_HD void camera_ray(const mrt_camera& c, float u, float v, float du, float dv, f3* o, f3* d) {
    const f3 pos = ld3(c.position), dir = ld3(c.direction), right = ld3(c.right), up = ld3(c.up);
    if (c.kind == MRT_CAMERA_PERSPECTIVE) {
        const float tvr = c.fov_h * (u - 0.5f);
        const float rf = fast_arctan(tvr) + du;
        const f3 rr = muls(right, rf);
        const float tvu = c.fov_v * (0.5f - v);
        const float uf = fast_arctan(tvu) + dv;
        const f3 uu = muls(up, uf);
        const f3 dest = add(add(add(pos, dir), rr), uu);
        *o = pos;
        *d = normalize(sub(dest, pos));
    } else {
        const float rf = (u - 0.5f) * c.fov_h;
        const f3 rr = add(muls(right, rf), muls(right, du));
        const float uf = (0.5f - v) * c.fov_v;
        const f3 uu = add(muls(up, uf), muls(up, dv));
        *o = add(add(pos, rr), uu);
        *d = dir;
    }
}

__device__ __forceinline__ MatV load_mat(const float4* m, uint32_t i) {
    const float4 a = m[4u * i + 0u], b = m[4u * i + 1u], c = m[4u * i + 2u], d = m[4u * i + 3u];
    MatV r;
    r.le = f3{a.x, a.y, a.z};
    r.kd = f3{a.w, b.x, b.y};
    r.ks = f3{b.z, b.w, c.x};
    r.kt = f3{c.y, c.z, c.w};
    r.ior = d.x;
    return r;
}

// What shade() needs from an Intersection (Intersection.hpp:15-21).
struct Surf {
    f3 point, normal, color;
    uint32_t excl;      // Intersection::primitive_ as an id (NONE for spheres)
    int32_t mat;        // material index, -1 = none (kD from the colour)
};

__device__ __forceinline__ f3 interp(f3 a, f3 b, f3 c, float w, float u, float v) {
    return add(add(muls(a, w), muls(b, u)), muls(c, v));
}

// Intersection fields of the hit (Triangle.cpp:89-101, Sphere.cpp:45-50,
// Plane.cpp:39-44, AreaLight.cpp:37-43) and Shader.cpp:153-155's material.
__device__ __forceinline__ Surf surface(const SceneDev& s, const RayT& r, const HitT& h) {
    Surf sf;
    const uint32_t cat = MRT_ID_CAT(h.id), idx = MRT_ID_INDEX(h.id);
    sf.point = add(r.o, muls(r.d, h.t));
    sf.excl = h.id;
    sf.color = f3{1.0f, 1.0f, 1.0f};
    if (cat == MRT_CAT_TRIANGLE || cat == MRT_CAT_LIGHT) {
        const float4* sh = (cat == MRT_CAT_TRIANGLE ? s.tri_shade : s.light_shade) + 5u * idx;
        const float4 a = sh[0], b = sh[1], c = sh[2], d = sh[3], e = sh[4];
        const float w = (1.0f - h.u) - h.v;
        sf.normal = interp(f3{a.x, a.y, a.z}, f3{a.w, b.x, b.y}, f3{b.z, b.w, c.x}, w, h.u, h.v);
        sf.color = interp(f3{c.y, c.z, c.w}, f3{d.x, d.y, d.z}, f3{d.w, e.x, e.y}, w, h.u, h.v);
        sf.mat = cat == MRT_CAT_LIGHT ? static_cast<int32_t>(s.light_info[idx].material) : __float_as_int(e.z);
    } else if (cat == MRT_CAT_SPHERE) {
        const float4 c = s.sph[idx];
        sf.normal = normalize(sub(sf.point, f3{c.x, c.y, c.z}));
        sf.mat = s.sph_mat[idx];
        sf.excl = MRT_ID_NONE;
    } else {
        const float4 n = s.pla[2u * idx];
        sf.normal = f3{n.x, n.y, n.z};
        sf.mat = __float_as_int(n.w);
    }
    return sf;
}

struct DrawCtx {
    uint32_t pixel, sample, branch;
};

__device__ __forceinline__ float draw_at(const RenderParams& p, const mrt_sampler& sm, const DrawCtx& c,
                                         uint32_t slot, uint32_t ls, int32_t depth) {
    return draw(sm, c.pixel, c.sample, draw_dim(slot, ls, static_cast<uint32_t>(depth), c.branch), p.spp);
}

// Shader::getCosineSampleHemisphere, Shader.cpp:193-223; cos and sin in
// double rounded to float (the oracle does the same).
__device__ __forceinline__ f3 cosine_hemisphere(const RenderParams& p, f3 n, const DrawCtx& c, int32_t depth) {
    const float r1 = draw_at(p, p.s_hemi, c, kSlotHemi1, 0, depth);
    const float r2 = draw_at(p, p.s_hemi, c, kSlotHemi2, 0, depth);
    const float phi = 6.28318530717958647692f * r1;
    const float cosTheta = sqrtf(r2);
    f3 u = fabsf(n.x) > 0.1f ? f3{0.0f, 1.0f, 0.0f} : f3{1.0f, 0.0f, 0.0f};
    u = normalize(cross(u, n));
    const f3 v = cross(n, u);
    const float cphi = static_cast<float>(cos(static_cast<double>(phi)));
    const float sphi = static_cast<float>(sin(static_cast<double>(phi)));
    const f3 dir = add(add(muls(u, cphi * cosTheta), muls(v, sphi * cosTheta)), muls(n, sqrtf(1.0f - r2)));
    return normalize(dir);
}

// The direct-light loop of Whitted.cpp:53-82 / PathTracer.cpp:71-103 with
// Shader::getLight (Shader.cpp:225-236) and Light::getPosition
// (AreaLight.cpp:22-31, PointLight.cpp:17-19).
// One light sample i of the direct-light loop: chooses the light, samples a
// point on it and returns true with the shadow-ray direction, distance and
// Le * cos when the light is above the surface (cos > 0).
__device__ __forceinline__ bool light_sample(const SceneDev& s, const RenderParams& p, const Surf& sf, int32_t depth,
                                             const DrawCtx& c, uint32_t i, f3* dir, float* dist, f3* lc) {
    const float rl = draw_at(p, p.s_light, c, kSlotLight, i, depth);
    const uint32_t li = static_cast<uint32_t>(floorf((rl * static_cast<float>(s.n_lights)) * 0.99999f));
    const LightInfo info = s.light_info[li];
    f3 lp;
    if (info.kind == MRT_LIGHT_AREA) {
        float R = draw_at(p, info.sampler, c, kSlotLightR, i, depth);
        float S = draw_at(p, info.sampler, c, kSlotLightS, i, depth);
        if (R + S >= 1.0f) {
            R = 1.0f - R;
            S = 1.0f - S;
        }
        const TriG g = load_tri(s.light_geo, li);
        lp = add(add(g.a, muls(g.ab, R)), muls(g.ac, S));
    } else {
        lp = f3{info.pos[0], info.pos[1], info.pos[2]};
    }
    f3 toLight = sub(lp, sf.point);
    *dist = length(toLight);
    toLight = normalize(toLight);
    const float cosNL = dot(sf.normal, toLight);
    if (!(cosNL > 0.0f)) return false;
    *dir = toLight;
    *lc = muls(load_mat(s.mat, info.material).le, cosNL);
    return true;
}

template <bool GENERIC, bool COUNT>
__device__ f3 direct_light(const SceneDev& s, const RenderParams& p, const Surf& sf, int32_t depth,
                           const DrawCtx& c, Stack& st, RayCounts& rc) {
    f3 L = f3{0.0f, 0.0f, 0.0f};
    for (uint32_t i = 0; i < p.samples_light; ++i) {
        f3 dir, lc;
        float dist;
        if (light_sample(s, p, sf, depth, c, i, &dir, &dist, &lc)) {
            ++rc.shadow;
            const RayT sr = make_ray(sf.point, dir, sf.excl);
            if (!any_hit<GENERIC, COUNT>(s, sr, dist, st, rc.ts)) L = add(L, lc);
        }
    }
    return L;
}

// Material of a hit (PathTracer.cpp:33-52): false with *le set when the hit
// is an emitter.
struct Shading {
    f3 kd, ks, kt, le;
    float ior;
    bool emitter;
};
__device__ __forceinline__ Shading shading_of(const SceneDev& s, const Surf& sf) {
    Shading m;
    m.ks = m.kt = m.le = f3{0.0f, 0.0f, 0.0f};
    m.ior = 1.0f;
    m.emitter = false;
    if (sf.mat >= 0) {
        const MatV mv = load_mat(s.mat, static_cast<uint32_t>(sf.mat));
        m.le = mv.le;
        m.emitter = any_gt0(mv.le);
        m.kd = mv.kd;
        m.ks = mv.ks;
        m.kt = mv.kt;
        m.ior = mv.ior;
    } else {
        m.kd = sf.color;
    }
    return m;
}

// A pending reflection / transmission chain.
struct Pending {
    f3 o, d, w;
    uint32_t excl, branch;
    int32_t depth;     // depth of the child ray
    int32_t tag;       // depth of the vertex that spawned it
    uint32_t discard;  // traced (the reference traces it) but its value is dropped
};
constexpr int kMaxPending = 40;

// Radiance of one pixel-sample starting from `ray` at `depth` (the value of
// Shader::rayTrace(&rgb, ray)); *rootFlag receives its return value.
// BRANCHES = false compiles out reflection / transmission.
template <bool GENERIC, bool BRANCHES, bool COUNT>
__device__ f3 trace_path(const SceneDev& s, const RenderParams& p, f3 o, f3 d, uint32_t excl, int32_t depth,
                         uint32_t pixel, uint32_t sample, bool* rootFlag, Stack& st, RayCounts& rc, int firstKind,
                         uint32_t* pendOverflow) {
    f3 pix = f3{0.0f, 0.0f, 0.0f};
    f3 T = f3{1.0f, 1.0f, 1.0f};
    DrawCtx c{pixel, sample, 0u};
    int kind = firstKind;  // 0 camera, 1 extension, 2 specular
    bool root = true;
    Pending pend[BRANCHES ? kMaxPending : 1];
    int npend = 0;
    bool discard = false;
    const bool whitted = p.shader == MRT_SHADER_WHITTED;
    while (true) {
        f3 accAll = f3{0.0f, 0.0f, 0.0f};
        f3 accPre = f3{0.0f, 0.0f, 0.0f};
        bool found = false, ends = false;
        int32_t dstar = 0;
        const int mark = npend;
        (void)mark;
        RayT r = make_ray(o, d, excl);
        while (depth <= p.depth_max) {
            const HitT h = closest_hit<GENERIC, COUNT>(s, r, kRayLengthMax, st, rc.ts);
            if (kind == 0) ++rc.camera;
            else if (kind == 1) ++rc.extension;
            else ++rc.specular;
            if (h.id == MRT_ID_NONE) break;
            const Surf sf = surface(s, r, h);
            const Shading m = shading_of(s, sf);
            if (m.emitter) {
                accAll = add(accAll, mul(T, m.le));
                ends = true;
                break;
            }
            const f3 kD = m.kd, kS = m.ks, kT = m.kt;
            const float ior = m.ior;
            const bool diffuse = any_gt0(kD);
            f3 Ld = f3{0.0f, 0.0f, 0.0f};
            if (diffuse && s.n_lights > 0u) {
                Ld = direct_light<GENERIC, COUNT>(s, p, sf, depth, c, st, rc);
                Ld = mul(Ld, kD);
                Ld = divs(Ld, static_cast<float>(p.samples_light));
            }
            if (BRANCHES && depth + 1 <= p.depth_max) {
                if (any_gt0(kT) && npend < kMaxPending) {
                    Pending& q = pend[npend++];
                    q.o = sf.point;
                    q.d = refract(r.d, sf.normal, 1.0f / ior);
                    q.w = mul(T, kT);
                    q.excl = sf.excl;
                    q.branch = c.branch * 3u + 2u;
                    q.depth = depth + 1;
                    q.tag = depth;
                    q.discard = discard ? 1u : 0u;
                } else if (any_gt0(kT)) {
                    *pendOverflow = 1u;
                }
                if (any_gt0(kS) && npend < kMaxPending) {
                    Pending& q = pend[npend++];
                    q.o = sf.point;
                    q.d = reflect(r.d, sf.normal);
                    q.w = mul(T, kS);
                    q.excl = sf.excl;
                    q.branch = c.branch * 3u + 1u;
                    q.depth = depth + 1;
                    q.tag = depth;
                    q.discard = discard ? 1u : 0u;
                } else if (any_gt0(kS)) {
                    *pendOverflow = 1u;
                }
            }
            if (whitted) {
                // Whitted.cpp:48-106: direct * kD / samplesLight, then the
                // ambient 0.1 * kD; reflection and transmission are queued.
                const f3 local = add(Ld, muls(kD, 0.1f));
                accAll = add(accAll, mul(T, local));
                break;
            }
            if (!diffuse) break;
            accAll = add(accAll, mul(T, Ld));
            if (!found && any_gt0(Ld)) {
                found = true;
                accPre = accAll;
                dstar = depth;
            }
            bool cont = depth <= p.depth_min;
            if (!cont) cont = draw_at(p, p.s_rr, c, kSlotRR, 0, depth) > 0.5f;
            if (!cont) break;
            if (depth + 1 > p.depth_max) break;
            const f3 nd = cosine_hemisphere(p, sf.normal, c, depth);
            T = mul(T, kD);
            if (depth > p.depth_min) T = divs(T, 0.5f * 0.5f);
            r = make_ray(sf.point, nd, sf.excl);
            ++depth;
            kind = 1;
        }
        const bool cut = ends && found;
        if (!discard) pix = add(pix, cut ? accPre : accAll);
        if (root) {
            *rootFlag = ends;
            root = false;
        }
        if (!BRANCHES) break;
        if (cut) {
            // Children queued below d* sit inside the LiD the reference zeroes:
            // it still traces them, so they run, but their value is dropped.
            for (int k = npend - 1; k >= mark && pend[k].tag > dstar; --k) pend[k].discard = 1u;
        }
        if (npend == 0) break;
        const Pending q = pend[--npend];
        o = q.o;
        d = q.d;
        excl = q.excl;
        depth = q.depth;
        T = q.w;
        c.branch = q.branch;
        discard = q.discard != 0u;
        kind = 2;
    }
    return pix;
}

__device__ __forceinline__ bool simple_shader(int32_t shader) {
    return shader == MRT_SHADER_NOSHADOWS || shader == MRT_SHADER_DEPTHMAP || shader == MRT_SHADER_DIFFUSE;
}

// Shader::rayTrace (Shader.cpp:124-160) of the one-hit shaders: NoShadows
// (NoShadows.cpp:19-64: emitters return Le; else the unoccluded direct light
// Le * cos over the light samples, times kD over samplesLight, plus 0.1 kD),
// DepthMap (DepthMap.cpp:19-26: grey max((maxDist - t) / maxDist, 0) with
// maxDist = 2 |maxPoint - origin|) and DiffuseMaterial
// (DiffuseMaterial.cpp:16-38: the first non-zero of kD, kS, kT, Le, or the
// primitive colour without a material).  A miss leaves the pixel black.
template <bool GENERIC, bool COUNT>
__device__ f3 trace_simple(const SceneDev& s, const RenderParams& p, f3 o, f3 d, uint32_t excl, int32_t depth,
                           uint32_t pixel, uint32_t sample, bool* flag, Stack& st, RayCounts& rc, int kind) {
    const RayT r = make_ray(o, d, excl);
    const HitT h = closest_hit<GENERIC, COUNT>(s, r, kRayLengthMax, st, rc.ts);
    if (kind == 0) ++rc.camera;
    else if (kind == 1) ++rc.extension;
    else ++rc.specular;
    *flag = false;
    f3 rgb = f3{0.0f, 0.0f, 0.0f};
    if (h.id == MRT_ID_NONE) return rgb;
    const Surf sf = surface(s, r, h);
    if (p.shader == MRT_SHADER_DEPTHMAP) {
        const float maxDist = length(sub(f3{p.max_point[0], p.max_point[1], p.max_point[2]}, o)) * 2.0f;
        const float v = fmaxf((maxDist - h.t) / maxDist, 0.0f);
        return f3{v, v, v};
    }
    if (p.shader == MRT_SHADER_DIFFUSE) {
        if (sf.mat < 0) return sf.color;
        const MatV mv = load_mat(s.mat, static_cast<uint32_t>(sf.mat));
        if (any_gt0(mv.kd)) return mv.kd;
        if (any_gt0(mv.ks)) return mv.ks;
        if (any_gt0(mv.kt)) return mv.kt;
        if (any_gt0(mv.le)) return mv.le;
        return rgb;
    }
    const Shading m = shading_of(s, sf);
    if (m.emitter) {
        *flag = true;
        return m.le;
    }
    if (any_gt0(m.kd) && s.n_lights > 0u) {
        const DrawCtx c{pixel, sample, 0u};
        for (uint32_t i = 0; i < p.samples_light; ++i) {
            f3 dir, lc;
            float dist;
            if (light_sample(s, p, sf, depth, c, i, &dir, &dist, &lc)) rgb = add(rgb, lc);
        }
        rgb = mul(rgb, m.kd);
        rgb = divs(rgb, static_cast<float>(p.samples_light));
    }
    return add(rgb, muls(m.kd, 0.1f));
}

}  // namespace mrt
