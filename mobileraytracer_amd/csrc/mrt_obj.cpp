// OBJ / MTL scene ingestion with the reference loader's semantics.
//
// The reference reads files with tinyobjloader (an un-vendored dependency
// with no pinned version) and converts its arrays in OBJLoader::fillScene
// (app/Components/ObjectLoaders/OBJLoader.cpp:66-250).  This parser reads the
// subset of the formats that conversion uses and applies the same rules:
//   OBJ  v x y z [r g b] (colour 1 when absent), vn, f with v, v/t, v//n,
//        v/t/n corners (1-based or negative relative indices), polygons split
//        into the fan (0, k-1, k) (tinyobjloader's triangulation), usemtl;
//        everything else ignored;
//   MTL  newmtl, Kd, Ks, Kt / Tf (transmittance), Ke (emission), Ni (ior),
//        d (dissolve), Tr (dissolve = 1 - Tr); tinyobjloader's defaults
//        (colours 0, ior 1, dissolve 1);
//   conversion (OBJLoader.cpp:82-241): X negated on vertices and normals;
//        normals normalised and used only when all three are non-zero; the
//        material's specular halved, transmittance times (1 - dissolve),
//        emission divided by its largest component when that exceeds 1; a
//        face with any emission > 0 becomes an AreaLight of that material
//        (triangle with material -1, face normal, zero colours); other faces
//        take the first equal material (Material.cpp:21-29: every component
//        within Epsilon) or append theirs; faces without a material keep -1.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "mrt_internal.hpp"

using namespace mrt;

namespace {

struct ObjMaterial {
    float kd[3] = {0, 0, 0}, ks[3] = {0, 0, 0}, kt[3] = {0, 0, 0}, ke[3] = {0, 0, 0};
    float ior = 1.0f, dissolve = 1.0f;
};

struct Corner {
    int32_t v, n;  // 0-based vertex and normal index, n = -1 when absent
};

bool read3(std::istringstream& in, float* out) {
    for (int k = 0; k < 3; ++k)
        if (!(in >> out[k])) return false;
    return true;
}

// 1-based or negative (relative to the current count) OBJ index -> 0-based;
// -1 when out of range.
int32_t resolve_index(long idx, size_t count) {
    long r = idx > 0 ? idx - 1 : static_cast<long>(count) + idx;
    return (idx == 0 || r < 0 || r >= static_cast<long>(count)) ? -1 : static_cast<int32_t>(r);
}

int parse_mtl(const char* path, std::vector<ObjMaterial>* mats, std::unordered_map<std::string, int32_t>* ids) {
    std::ifstream f(path);
    if (!f) return fail(MRT_E_INVALID, std::string("cannot open MTL file ") + path);
    std::string line;
    ObjMaterial* cur = nullptr;
    while (std::getline(f, line)) {
        std::istringstream in(line);
        std::string key;
        if (!(in >> key) || key[0] == '#') continue;
        if (key == "newmtl") {
            std::string name;
            in >> name;
            (*ids)[name] = static_cast<int32_t>(mats->size());
            mats->push_back(ObjMaterial{});
            cur = &mats->back();
            continue;
        }
        if (!cur) continue;
        float v[3];
        if (key == "Kd" && read3(in, v)) std::memcpy(cur->kd, v, sizeof v);
        else if (key == "Ks" && read3(in, v)) std::memcpy(cur->ks, v, sizeof v);
        else if ((key == "Kt" || key == "Tf") && read3(in, v)) std::memcpy(cur->kt, v, sizeof v);
        else if (key == "Ke" && read3(in, v)) std::memcpy(cur->ke, v, sizeof v);
        else if (key == "Ni") in >> cur->ior;
        else if (key == "d") in >> cur->dissolve;
        else if (key == "Tr") {
            float tr = 0.0f;
            if (in >> tr) cur->dissolve = 1.0f - tr;
        }
    }
    return MRT_OK;
}

// Material.cpp:21-29.
bool same_material(const mrt_material& a, const mrt_material& b) {
    for (int k = 0; k < 3; ++k) {
        if (!(std::fabs(a.le[k] - b.le[k]) <= kEpsilon) || !(std::fabs(a.kd[k] - b.kd[k]) <= kEpsilon) ||
            !(std::fabs(a.ks[k] - b.ks[k]) <= kEpsilon) || !(std::fabs(a.kt[k] - b.kt[k]) <= kEpsilon))
            return false;
    }
    return std::fabs(a.ior - b.ior) < kEpsilon;
}

void make_tri(const f3& a, const f3& b, const f3& c, int32_t mat, const f3* normals, const f3* colors,
              mrt_triangle* out) {
    std::memset(out, 0, sizeof(*out));
    const f3 ac = sub(c, a), ab = sub(b, a);
    st3(out->ac, ac);
    st3(out->ab, ab);
    st3(out->a, a);
    if (normals) {
        st3(out->na, normals[0]);
        st3(out->nb, normals[1]);
        st3(out->nc, normals[2]);
    } else {
        const f3 n = normalize(cross(ac, ab));  // Triangle.cpp:12-22
        st3(out->na, n);
        st3(out->nb, n);
        st3(out->nc, n);
    }
    if (colors) {
        st3(out->ca, colors[0]);
        st3(out->cb, colors[1]);
        st3(out->cc, colors[2]);
    }
    out->material_id = mat;
}

}  // namespace

struct mrt_obj {
    std::vector<mrt_triangle> tris;
    std::vector<mrt_material> mats;
    std::vector<mrt_material> light_radiance;
    std::vector<mrt_triangle> light_tris;
    uint32_t faces = 0, skipped = 0;
};

namespace {

int load_obj(const char* obj_path, const char* mtl_path, mrt_obj* o) {
    std::vector<ObjMaterial> omats;
    std::unordered_map<std::string, int32_t> ids;
    int rc;
    if (mtl_path && mtl_path[0] && (rc = parse_mtl(mtl_path, &omats, &ids)) != MRT_OK) return rc;
    std::ifstream f(obj_path);
    if (!f) return fail(MRT_E_INVALID, std::string("cannot open OBJ file ") + obj_path);
    std::vector<f3> v, vc, vn;
    std::vector<Corner> corners;   // 3 per triangle
    std::vector<int32_t> faceMat;  // per triangle
    int32_t curMat = -1;
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream in(line);
        std::string key;
        if (!(in >> key) || key[0] == '#') continue;
        if (key == "v") {
            float x[7];
            int n = 0;
            while (n < 7 && (in >> x[n])) ++n;
            if (n < 3) return fail(MRT_E_INVALID, "OBJ vertex with fewer than 3 coordinates");
            v.push_back(f3{x[0], x[1], x[2]});
            vc.push_back(n >= 6 ? f3{x[3], x[4], x[5]} : f3{1.0f, 1.0f, 1.0f});
        } else if (key == "vn") {
            float x[3];
            if (!read3(in, x)) return fail(MRT_E_INVALID, "OBJ normal with fewer than 3 coordinates");
            vn.push_back(f3{x[0], x[1], x[2]});
        } else if (key == "usemtl") {
            std::string name;
            in >> name;
            const auto it = ids.find(name);
            curMat = it == ids.end() ? -1 : it->second;
        } else if (key == "f") {
            std::vector<Corner> poly;
            std::string tok;
            while (in >> tok) {
                long vi = 0, ni = 0;
                bool hasN = false;
                const size_t s1 = tok.find('/');
                vi = std::strtol(tok.c_str(), nullptr, 10);
                if (s1 != std::string::npos) {
                    const size_t s2 = tok.find('/', s1 + 1);
                    if (s2 != std::string::npos && s2 + 1 < tok.size()) {
                        ni = std::strtol(tok.c_str() + s2 + 1, nullptr, 10);
                        hasN = true;
                    }
                }
                const int32_t a = resolve_index(vi, v.size());
                if (a < 0) return fail(MRT_E_INVALID, "OBJ face with a vertex index out of range: " + tok);
                poly.push_back(Corner{a, hasN ? resolve_index(ni, vn.size()) : -1});
            }
            ++o->faces;
            if (poly.size() < 3) {
                ++o->skipped;
                continue;
            }
            for (size_t k = 2; k < poly.size(); ++k) {
                corners.push_back(poly[0]);
                corners.push_back(poly[k - 1]);
                corners.push_back(poly[k]);
                faceMat.push_back(curMat);
            }
        }
    }
    // OBJLoader::fillScene, OBJLoader.cpp:66-250.
    const size_t nt = faceMat.size();
    for (size_t t = 0; t < nt; ++t) {
        f3 p[3], n[3], c[3];
        bool hasNormal = !vn.empty();
        for (int k = 0; k < 3; ++k) {
            const Corner& q = corners[3 * t + k];
            p[k] = f3{-v[q.v].x, v[q.v].y, v[q.v].z};
            c[k] = vc[q.v];
            if (q.n >= 0) n[k] = normalize(f3{-vn[q.n].x, vn[q.n].y, vn[q.n].z});
            else n[k] = f3{0.0f, 0.0f, 0.0f};
            if (!(length(n[k]) > 0.0f)) hasNormal = false;
        }
        const int32_t m = faceMat[t];
        if (m < 0) {
            mrt_triangle tri;
            make_tri(p[0], p[1], p[2], -1, hasNormal ? n : nullptr, c, &tri);
            o->tris.push_back(tri);
            continue;
        }
        const ObjMaterial& om = omats[static_cast<size_t>(m)];
        mrt_material mat{};
        float e[3] = {om.ke[0], om.ke[1], om.ke[2]};
        const float mx = std::max(std::max(e[0], e[1]), e[2]);
        if (mx > 1.0f) {
            e[0] /= mx;
            e[1] /= mx;
            e[2] /= mx;
        }
        for (int k = 0; k < 3; ++k) {
            mat.kd[k] = om.kd[k];
            mat.ks[k] = om.ks[k] / 2.0f;
            mat.kt[k] = om.kt[k] * (1.0f - om.dissolve);
            mat.le[k] = e[k];
        }
        mat.ior = om.ior;
        if (e[0] > 0.0f || e[1] > 0.0f || e[2] > 0.0f) {
            mrt_triangle tri;
            make_tri(p[0], p[1], p[2], -1, nullptr, nullptr, &tri);
            o->light_radiance.push_back(mat);
            o->light_tris.push_back(tri);
            continue;
        }
        int32_t idx = -1;
        for (size_t k = 0; k < o->mats.size(); ++k)
            if (same_material(o->mats[k], mat)) {
                idx = static_cast<int32_t>(k);
                break;
            }
        if (idx < 0) {
            idx = static_cast<int32_t>(o->mats.size());
            o->mats.push_back(mat);
        }
        mrt_triangle tri;
        make_tri(p[0], p[1], p[2], idx, hasNormal ? n : nullptr, c, &tri);
        o->tris.push_back(tri);
    }
    return MRT_OK;
}

}  // namespace

extern "C" {

MRT_API int mrt_obj_load(const char* obj_path, const char* mtl_path, mrt_obj** out) {
    if (!obj_path || !out) return fail(MRT_E_INVALID, "null argument");
    *out = nullptr;
    auto* o = new mrt_obj;
    const int rc = load_obj(obj_path, mtl_path, o);
    if (rc != MRT_OK) {
        delete o;
        return rc;
    }
    *out = o;
    return MRT_OK;
}

MRT_API int mrt_obj_counts(const mrt_obj* o, uint32_t* n_triangles, uint32_t* n_materials, uint32_t* n_lights,
                           uint32_t* n_faces_skipped) {
    if (!o) return fail(MRT_E_INVALID, "null OBJ");
    if (n_triangles) *n_triangles = static_cast<uint32_t>(o->tris.size());
    if (n_materials) *n_materials = static_cast<uint32_t>(o->mats.size());
    if (n_lights) *n_lights = static_cast<uint32_t>(o->light_tris.size());
    if (n_faces_skipped) *n_faces_skipped = o->skipped;
    return MRT_OK;
}

MRT_API int mrt_obj_fill(const mrt_obj* o, mrt_triangle* tris, mrt_material* mats, mrt_light* lights,
                         const mrt_sampler* light_sampler) {
    if (!o) return fail(MRT_E_INVALID, "null OBJ");
    if ((!o->tris.empty() && !tris) || (!o->mats.empty() && !mats) || (!o->light_tris.empty() && !lights))
        return fail(MRT_E_INVALID, "null output array with a non-zero count");
    if (!o->tris.empty()) std::memcpy(tris, o->tris.data(), o->tris.size() * sizeof(mrt_triangle));
    if (!o->mats.empty()) std::memcpy(mats, o->mats.data(), o->mats.size() * sizeof(mrt_material));
    for (size_t k = 0; k < o->light_tris.size(); ++k) {
        mrt_light& l = lights[k];
        std::memset(&l, 0, sizeof(l));
        l.kind = MRT_LIGHT_AREA;
        l.radiance = o->light_radiance[k];
        l.triangle = o->light_tris[k];
        // A sampler per light, as the reference's lambda() gives each
        // AreaLight its own generator: the seed advances per light.
        l.sampler = light_sampler ? *light_sampler : mrt_sampler{MRT_SAMPLER_RANDOM, 0.0f, 0u, 0u};
        l.sampler.seed += static_cast<uint32_t>(k);
    }
    return MRT_OK;
}

MRT_API int mrt_obj_destroy(mrt_obj* o) {
    delete o;
    return MRT_OK;
}

}  // extern "C"
