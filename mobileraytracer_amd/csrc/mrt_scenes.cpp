// Host construction helpers (Triangle / Camera constructors, the tile table)
// and the seeded synthetic scenes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "mrt_build.hpp"
#include "mrt_sampling.hpp"

namespace {

using namespace mrt;

// splitmix64 (Steele, Lea, Flood 2014) as the generators' random source.
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    float uniform() { return static_cast<float>(next() >> 40) * 5.9604644775390625e-08f; }
    float range(float a, float b) { return a + (b - a) * uniform(); }
};

void make_triangle(f3 a, f3 b, f3 c, int32_t mat, const float* normals9, const float* colors9, mrt_triangle* out) {
    std::memset(out, 0, sizeof(*out));
    const f3 ac = sub(c, a), ab = sub(b, a);
    st3(out->ac, ac);
    st3(out->ab, ab);
    st3(out->a, a);
    if (normals9) {
        std::memcpy(out->na, normals9, 3 * sizeof(float));
        std::memcpy(out->nb, normals9 + 3, 3 * sizeof(float));
        std::memcpy(out->nc, normals9 + 6, 3 * sizeof(float));
    } else {
        const f3 n = normalize(cross(ac, ab));  // Triangle.cpp:12-22
        st3(out->na, n);
        st3(out->nb, n);
        st3(out->nc, n);
    }
    if (colors9) {
        std::memcpy(out->ca, colors9, 3 * sizeof(float));
        std::memcpy(out->cb, colors9 + 3, 3 * sizeof(float));
        std::memcpy(out->cc, colors9 + 6, 3 * sizeof(float));
    }
    out->material_id = mat;
}

// Two triangles of a quad p0 p1 p2 p3 (in order around the rim) whose face
// normal normalize(cross(AC, AB)) points along `facing`.
void quad(f3 p0, f3 p1, f3 p2, f3 p3, f3 facing, int32_t mat, mrt_triangle* out) {
    const f3 n = cross(sub(p2, p0), sub(p1, p0));
    if (dot(n, facing) < 0.0f) std::swap(p1, p3);
    make_triangle(p0, p1, p2, mat, nullptr, nullptr, out);
    make_triangle(p0, p2, p3, mat, nullptr, nullptr, out + 1);
}

// Axis-aligned box [lo, hi], faces wound outwards.
void box(f3 lo, f3 hi, int32_t mat, mrt_triangle* out) {
    const f3 c000{lo.x, lo.y, lo.z}, c100{hi.x, lo.y, lo.z}, c110{hi.x, hi.y, lo.z}, c010{lo.x, hi.y, lo.z};
    const f3 c001{lo.x, lo.y, hi.z}, c101{hi.x, lo.y, hi.z}, c111{hi.x, hi.y, hi.z}, c011{lo.x, hi.y, hi.z};
    quad(c000, c100, c110, c010, f3{0, 0, -1}, mat, out + 0);   // front  z = lo
    quad(c001, c011, c111, c101, f3{0, 0, 1}, mat, out + 2);    // back   z = hi
    quad(c000, c010, c011, c001, f3{-1, 0, 0}, mat, out + 4);   // left   x = lo
    quad(c100, c101, c111, c110, f3{1, 0, 0}, mat, out + 6);    // right  x = hi
    quad(c000, c001, c101, c100, f3{0, -1, 0}, mat, out + 8);   // bottom y = lo
    quad(c010, c110, c111, c011, f3{0, 1, 0}, mat, out + 10);   // top    y = hi
}

void set_mat(mrt_material* m, f3 kd, f3 le) {
    std::memset(m, 0, sizeof(*m));
    st3(m->kd, kd);
    st3(m->le, le);
    m->ior = 1.0f;
}

void area_light(f3 a, f3 b, f3 c, float le, int sampler_kind, uint32_t seed, mrt_light* out) {
    std::memset(out, 0, sizeof(*out));
    out->kind = MRT_LIGHT_AREA;
    set_mat(&out->radiance, f3{0, 0, 0}, f3{le, le, le});
    make_triangle(a, b, c, -1, nullptr, nullptr, &out->triangle);
    out->sampler.kind = sampler_kind;
    out->sampler.value = 0.5f;
    out->sampler.seed = seed;
    out->sampler.domain_size = 0;
}

// 30-bit Morton code of a point in the unit cube.
uint32_t morton30(f3 c) {
    auto spread = [](uint32_t v) {
        v &= 0x3FFu;
        v = (v | (v << 16)) & 0x030000FFu;
        v = (v | (v << 8)) & 0x0300F00Fu;
        v = (v | (v << 4)) & 0x030C30C3u;
        v = (v | (v << 2)) & 0x09249249u;
        return v;
    };
    auto q = [](float v) { return static_cast<uint32_t>(std::min(1023.0f, std::max(0.0f, v * 1024.0f))); };
    return (spread(q(c.x)) << 2) | (spread(q(c.y)) << 1) | spread(q(c.z));
}

}  // namespace

extern "C" {

MRT_API int mrt_triangle_make(const float a[3], const float b[3], const float c[3], int32_t material_id,
                              const float* normals9, const float* colors9, mrt_triangle* out) {
    if (!a || !b || !c || !out) return MRT_E_INVALID;
    make_triangle(ld3(a), ld3(b), ld3(c), material_id, normals9, colors9, out);
    return MRT_OK;
}

// Camera::Camera (Camera.cpp:47-54), Perspective ctor (Perspective.cpp:12-18,
// degToRad of Camera.cpp:72-75), Orthographic ctor (Orthographic.cpp:11-17).
MRT_API int mrt_camera_make(int kind, const float position[3], const float look_at[3], const float up[3],
                            float a, float b, mrt_camera* out) {
    if (!position || !look_at || !up || !out) return MRT_E_INVALID;
    if (kind != MRT_CAMERA_PERSPECTIVE && kind != MRT_CAMERA_ORTHOGRAPHIC) return MRT_E_INVALID;
    const f3 P = ld3(position), L = ld3(look_at), U = ld3(up);
    const f3 dir = normalize(sub(L, P));
    const f3 right = cross(U, dir);
    const f3 upv = cross(dir, right);
    std::memset(out, 0, sizeof(*out));
    out->kind = kind;
    st3(out->position, P);
    st3(out->direction, dir);
    st3(out->right, right);
    st3(out->up, upv);
    if (kind == MRT_CAMERA_PERSPECTIVE) {
        const float pi = 3.14159265358979323846264338327950288f;
        out->fov_h = (a * pi) / 180.0f;
        out->fov_v = (b * pi) / 180.0f;
    } else {
        out->fov_h = a / 2.0f;
        out->fov_v = b / 2.0f;
    }
    return MRT_OK;
}

MRT_API int mrt_camera_aabb(const mrt_camera* c, float out[6]) {
    if (!c || !out) return MRT_E_INVALID;
    const f3 P = ld3(c->position);
    f3 mn = P, mx = P;
    if (c->kind == MRT_CAMERA_ORTHOGRAPHIC) {
        const f3 R = ld3(c->right), U = ld3(c->up);
        const float sh = c->fov_h, sv = c->fov_v;  // sizeH_, sizeV_
        mn = add(add(add(add(P, muls(muls(R, 0.0f - 0.5f), sh)), muls(R, -0.5f)), muls(muls(U, 0.5f - 0.0f), sv)),
                 muls(U, -0.5f));
        mx = add(add(add(add(P, muls(muls(R, 1.0f - 0.5f), sh)), muls(R, 0.5f)), muls(muls(U, 0.5f - 1.0f), sv)),
                 muls(U, 0.5f));
    } else if (c->kind != MRT_CAMERA_PERSPECTIVE) {
        return MRT_E_INVALID;
    }
    st3(out, mn);
    st3(out + 3, mx);
    return MRT_OK;
}

// FillThings (Camera.cpp:16-26) with a seeded shuffle in place of the
// /dev/urandom-seeded std::shuffle.
MRT_API int mrt_tile_values(uint32_t seed, float out[256]) {
    if (!out) return MRT_E_INVALID;
    for (uint32_t i = 0; i < kBlocks; ++i) out[i] = halton(i, 2u);
    if (seed != 0u) {
        SplitMix rng{seed};
        for (uint32_t i = kBlocks - 1; i > 0; --i) {
            const uint32_t j = static_cast<uint32_t>(rng.next() % (i + 1));
            std::swap(out[i], out[j]);
        }
    }
    return MRT_OK;
}

MRT_API int mrt_triangle_intersect_box(const mrt_triangle* tri, const float box_min[3], const float box_max[3]) {
    if (!tri || !box_min || !box_max) return MRT_E_INVALID;
    return triangle_intersect_box(*tri, Box{ld3(box_min), ld3(box_max)}) ? 1 : 0;
}

MRT_API int mrt_primitive_intersect_box(int category, const void* prim, const float box_min[3],
                                        const float box_max[3]) {
    if (!prim || !box_min || !box_max) return MRT_E_INVALID;
    const Box box{ld3(box_min), ld3(box_max)};
    switch (category) {
        case MRT_CAT_TRIANGLE:
        case MRT_CAT_LIGHT: return triangle_intersect_box(*static_cast<const mrt_triangle*>(prim), box) ? 1 : 0;
        case MRT_CAT_SPHERE: return sphere_intersect_box(*static_cast<const mrt_sphere*>(prim), box) ? 1 : 0;
        case MRT_CAT_PLANE: return plane_intersect_box(*static_cast<const mrt_plane*>(prim), box) ? 1 : 0;
        default: return MRT_E_INVALID;
    }
}

MRT_API int mrt_primitive_aabbs(int category, const void* prims, uint32_t n, float* boxes6) {
    if ((n && (!prims || !boxes6)) || category < MRT_CAT_PLANE || category > MRT_CAT_LIGHT) return MRT_E_INVALID;
    for (uint32_t i = 0; i < n; ++i) {
        Box b;
        if (category == MRT_CAT_TRIANGLE || category == MRT_CAT_LIGHT)
            b = triangle_box(static_cast<const mrt_triangle*>(prims)[i]);
        else if (category == MRT_CAT_SPHERE)
            b = sphere_box(static_cast<const mrt_sphere*>(prims)[i]);
        else
            b = plane_box(static_cast<const mrt_plane*>(prims)[i]);
        st3(boxes6 + 6 * static_cast<size_t>(i), b.mn);
        st3(boxes6 + 6 * static_cast<size_t>(i) + 3, b.mx);
    }
    return MRT_OK;
}

MRT_API int mrt_bvh_build(const float* boxes6, uint32_t n, uint32_t leaf_size, mrt_bvh_node* nodes, uint32_t capacity,
                          uint32_t* count) {
    if ((n && !boxes6) || !count || leaf_size == 0) return MRT_E_INVALID;
    std::vector<Box> boxes(n);
    for (uint32_t i = 0; i < n; ++i)
        boxes[i] = Box{ld3(boxes6 + 6 * static_cast<size_t>(i)), ld3(boxes6 + 6 * static_cast<size_t>(i) + 3)};
    BvhOut out;
    build_bvh_reference(boxes, leaf_size, &out);
    *count = static_cast<uint32_t>(out.ref.size());
    if (nodes) {
        for (uint32_t i = 0; i < *count && i < capacity; ++i) {
            st3(nodes[i].min, out.ref[i].box.mn);
            st3(nodes[i].max, out.ref[i].box.mx);
            nodes[i].index_offset = out.ref[i].offset;
            nodes[i].number_primitives = out.ref[i].count;
        }
    }
    return MRT_OK;
}

MRT_API int mrt_wide_build(const float* boxes6, const uint32_t* words, uint32_t n, float* nodes, uint32_t capacity,
                           uint32_t* count, uint32_t* depth) {
    if ((n && (!boxes6 || !words)) || !count || !depth) return MRT_E_INVALID;
    std::vector<WideSlot> units(n);
    for (uint32_t i = 0; i < n; ++i) {
        units[i].box = Box{ld3(boxes6 + 6 * static_cast<size_t>(i)), ld3(boxes6 + 6 * static_cast<size_t>(i) + 3)};
        units[i].ref = i;
        units[i].word = words[i];
        if (words[i] == 0u) return MRT_E_INVALID;
    }
    std::vector<float4> out;
    std::vector<WideSlot> top;
    build_wide_sah(units, 1u, 4, &out, &top, depth);
    std::vector<float4> all;
    put_wide_node(&all, 0u, top.data(), static_cast<int>(top.size()));
    all.insert(all.end(), out.begin(), out.end());
    *count = static_cast<uint32_t>(all.size() / 8);
    if (nodes) std::memcpy(nodes, all.data(), std::min<size_t>(all.size(), 8ull * capacity) * sizeof(float4));
    return MRT_OK;
}

// Procedural Cornell box: walls [-1,1]^3 without the front (10 triangles,
// wound inwards), two axis-aligned boxes (24 triangles, wound outwards), one
// 0.5 x 0.5 area-light quad at y = 0.99 split in two AreaLights with Le 0.9
// (the pattern of app/Scenes/Scenes.cpp:91-129), albedos U[0.1,0.9]^3 per
// object.  Light samplers are CONSTANT 0.5.
MRT_API int mrt_scene_cornell(uint32_t seed, mrt_triangle* tris, mrt_material* mats, mrt_light* lights) {
    if (!tris || !mats || !lights) return MRT_E_INVALID;
    SplitMix rng{seed};
    for (uint32_t m = 0; m < MRT_CORNELL_MATERIALS; ++m) {
        const f3 kd{rng.range(0.1f, 0.9f), rng.range(0.1f, 0.9f), rng.range(0.1f, 0.9f)};
        set_mat(&mats[m], kd, f3{0, 0, 0});
    }
    const float L = -1.0f, H = 1.0f;
    quad(f3{L, L, L}, f3{H, L, L}, f3{H, L, H}, f3{L, L, H}, f3{0, 1, 0}, 0, tris + 0);   // floor
    quad(f3{L, H, L}, f3{H, H, L}, f3{H, H, H}, f3{L, H, H}, f3{0, -1, 0}, 1, tris + 2);  // ceiling
    quad(f3{L, L, H}, f3{H, L, H}, f3{H, H, H}, f3{L, H, H}, f3{0, 0, -1}, 2, tris + 4);  // back
    quad(f3{L, L, L}, f3{L, L, H}, f3{L, H, H}, f3{L, H, L}, f3{1, 0, 0}, 3, tris + 6);   // left
    quad(f3{H, L, L}, f3{H, L, H}, f3{H, H, H}, f3{H, H, L}, f3{-1, 0, 0}, 4, tris + 8);  // right
    box(f3{-0.70f, -0.99f, 0.05f}, f3{-0.15f, 0.25f, 0.62f}, 5, tris + 10);               // tall box
    box(f3{0.10f, -0.99f, -0.55f}, f3{0.68f, -0.38f, 0.03f}, 6, tris + 22);               // short box
    const f3 a{-0.25f, 0.99f, -0.25f}, b{0.25f, 0.99f, -0.25f}, c{0.25f, 0.99f, 0.25f}, d{-0.25f, 0.99f, 0.25f};
    area_light(a, b, c, 0.9f, MRT_SAMPLER_CONSTANT, seed, &lights[0]);
    area_light(c, d, a, 0.9f, MRT_SAMPLER_CONSTANT, seed + 1u, &lights[1]);
    return MRT_OK;
}

// Random triangle soup: n triangles with centroids U[0,1]^3 and vertices
// centroid + U(-h, h)^3, emitted in Morton order of their centroids (the
// reference builder keeps the input order, BVH.hpp:87-155, so spatially
// sorted input keeps its trees shallow).  n_emissive triangles, chosen by a
// seeded partial shuffle, become AreaLights with radiance le and the given
// point-sampler kind; the others go to `tris` with material -1 and a vertex
// colour U(0.2, 0.8)^3 as their Lambertian albedo.
MRT_API int mrt_scene_soup(uint32_t seed, uint32_t n, uint32_t n_emissive, float le, float h, mrt_triangle* tris,
                           mrt_light* lights, int light_sampler_kind) {
    if (!tris || (n_emissive && !lights) || n_emissive > n) return MRT_E_INVALID;
    SplitMix rng{seed};
    struct Raw {
        f3 v[3];
        f3 albedo;
        uint32_t code;
    };
    std::vector<Raw> raw(n);
    for (uint32_t i = 0; i < n; ++i) {
        const f3 c{rng.uniform(), rng.uniform(), rng.uniform()};
        for (int k = 0; k < 3; ++k)
            raw[i].v[k] = f3{c.x + rng.range(-h, h), c.y + rng.range(-h, h), c.z + rng.range(-h, h)};
        raw[i].albedo = f3{rng.range(0.2f, 0.8f), rng.range(0.2f, 0.8f), rng.range(0.2f, 0.8f)};
        raw[i].code = morton30(c);
    }
    std::vector<uint8_t> emissive(n, 0);
    {
        std::vector<uint32_t> idx(n);
        std::iota(idx.begin(), idx.end(), 0u);
        for (uint32_t i = 0; i < n_emissive; ++i) {
            const uint32_t j = i + static_cast<uint32_t>(rng.next() % (n - i));
            std::swap(idx[i], idx[j]);
            emissive[idx[i]] = 1;
        }
    }
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return raw[a].code < raw[b].code; });
    uint32_t nt = 0, nl = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const Raw& r = raw[order[k]];
        if (emissive[order[k]]) {
            area_light(r.v[0], r.v[1], r.v[2], le, light_sampler_kind, seed * 2654435761u + nl, &lights[nl]);
            ++nl;
        } else {
            const float col[9] = {r.albedo.x, r.albedo.y, r.albedo.z, r.albedo.x, r.albedo.y,
                                  r.albedo.z, r.albedo.x, r.albedo.y, r.albedo.z};
            make_triangle(r.v[0], r.v[1], r.v[2], -1, nullptr, col, &tris[nt]);
            ++nt;
        }
    }
    return MRT_OK;
}

// Cluster of displaced, subdivided icospheres in the unit cube.  Each sphere
// is an icosahedron split `subdiv` times into four (20 * 4^subdiv
// triangles) with its vertices on the unit sphere, then scaled by a radius
// U(0.03, 0.12), displaced along the direction by displace * radius * a sum of
// four seeded sinusoids of the direction (so vertices shared by neighbouring
// triangles stay shared) and moved to a centre U(0.15, 0.85)^3; one albedo
// U(0.2, 0.8)^3 per sphere.  All triangles are then ordered by the Morton code
// of their centroid, as in mrt_scene_soup, and n_emissive of them (seeded
// choice) become area lights.
MRT_API int mrt_scene_icospheres(uint32_t seed, uint32_t n_spheres, uint32_t subdiv, float displace,
                                 uint32_t n_emissive, float le, mrt_triangle* tris, mrt_light* lights,
                                 int light_sampler_kind) {
    if (subdiv > 10 || n_spheres == 0) return MRT_E_INVALID;
    const uint64_t perSphere = 20ull << (2 * subdiv);
    const uint64_t total = perSphere * n_spheres;
    if (total >= (1ull << 30) || !tris || n_emissive > total || (n_emissive && !lights)) return MRT_E_INVALID;
    const uint32_t n = static_cast<uint32_t>(total);
    SplitMix rng{seed ^ 0x1c05ull};
    // Icosahedron: 12 vertices (0, +-1, +-t) cyclic, 20 faces.
    const float t = (1.0f + std::sqrt(5.0f)) * 0.5f;
    const f3 iv[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                       {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    const int faces[20][3] = {{0, 11, 5}, {0, 5, 1},  {0, 1, 7},   {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                              {11, 10, 2}, {10, 7, 6}, {7, 1, 8},  {3, 9, 4},  {3, 4, 2},   {3, 2, 6}, {3, 6, 8},
                              {3, 8, 9},  {4, 9, 5},  {2, 4, 11}, {6, 2, 10}, {8, 6, 7},   {9, 8, 1}};
    std::vector<f3> unit;  // 3 directions per triangle of one sphere
    unit.reserve(3 * perSphere);
    for (const auto& f : faces)
        for (int k = 0; k < 3; ++k) unit.push_back(normalize(iv[f[k]]));
    for (uint32_t level = 0; level < subdiv; ++level) {
        std::vector<f3> next;
        next.reserve(unit.size() * 4);
        for (size_t i = 0; i < unit.size(); i += 3) {
            const f3 a = unit[i], b = unit[i + 1], c = unit[i + 2];
            const f3 ab = normalize(add(a, b)), bc = normalize(add(b, c)), ca = normalize(add(c, a));
            const f3 tri[4][3] = {{a, ab, ca}, {ab, b, bc}, {ca, bc, c}, {ab, bc, ca}};
            for (const auto& q : tri)
                for (int k = 0; k < 3; ++k) next.push_back(q[k]);
        }
        unit.swap(next);
    }
    struct Raw {
        f3 v[3];
        f3 albedo;
        uint32_t code;
    };
    std::vector<Raw> raw(n);
    uint32_t w = 0;
    for (uint32_t sp = 0; sp < n_spheres; ++sp) {
        const f3 centre{rng.range(0.15f, 0.85f), rng.range(0.15f, 0.85f), rng.range(0.15f, 0.85f)};
        const float radius = rng.range(0.03f, 0.12f);
        const f3 albedo{rng.range(0.2f, 0.8f), rng.range(0.2f, 0.8f), rng.range(0.2f, 0.8f)};
        f3 wave[4];
        float freq[4], phase[4];
        for (int k = 0; k < 4; ++k) {
            wave[k] = normalize(f3{rng.range(-1, 1), rng.range(-1, 1), rng.range(-1, 1) + 1e-3f});
            freq[k] = rng.range(2.0f, 9.0f);
            phase[k] = rng.range(0.0f, 6.2831853f);
        }
        auto place = [&](f3 u) {
            float d = 0.0f;
            for (int k = 0; k < 4; ++k) d += 0.25f * std::sin(freq[k] * dot(u, wave[k]) + phase[k]);
            return add(centre, muls(u, radius * (1.0f + displace * d)));
        };
        for (size_t i = 0; i < unit.size(); i += 3, ++w) {
            for (int k = 0; k < 3; ++k) raw[w].v[k] = place(unit[i + k]);
            raw[w].albedo = albedo;
            const f3 c = muls(add(add(raw[w].v[0], raw[w].v[1]), raw[w].v[2]), 1.0f / 3.0f);
            raw[w].code = morton30(f3{std::min(std::max(c.x, 0.0f), 1.0f), std::min(std::max(c.y, 0.0f), 1.0f),
                                      std::min(std::max(c.z, 0.0f), 1.0f)});
        }
    }
    std::vector<uint8_t> emissive(n, 0);
    {
        std::vector<uint32_t> idx(n);
        std::iota(idx.begin(), idx.end(), 0u);
        for (uint32_t i = 0; i < n_emissive; ++i) {
            const uint32_t j = i + static_cast<uint32_t>(rng.next() % (n - i));
            std::swap(idx[i], idx[j]);
            emissive[idx[i]] = 1;
        }
    }
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return raw[a].code < raw[b].code; });
    uint32_t nt = 0, nl = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const Raw& r = raw[order[k]];
        if (emissive[order[k]]) {
            area_light(r.v[0], r.v[1], r.v[2], le, light_sampler_kind, seed * 2654435761u + nl, &lights[nl]);
            ++nl;
        } else {
            const float col[9] = {r.albedo.x, r.albedo.y, r.albedo.z, r.albedo.x, r.albedo.y,
                                  r.albedo.z, r.albedo.x, r.albedo.y, r.albedo.z};
            make_triangle(r.v[0], r.v[1], r.v[2], -1, nullptr, col, &tris[nt]);
            ++nt;
        }
    }
    return MRT_OK;
}

}  // extern "C"
