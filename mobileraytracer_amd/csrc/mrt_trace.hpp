// Device traversal: primitive tests, BVH / RegularGrid / Naive walks and the
// light scan, for closest hit (Shader::rayTrace's intersection step,
// app/MobileRT/Shader.cpp:124-156) and any hit (Shader::shadowTrace,
// Shader.cpp:162-187).
//
// Result parity with the reference:
//  * The primitive tests repeat the reference arithmetic operation for
//    operation (Triangle.cpp:57-136, Sphere.cpp:18-64, Plane.cpp:17-57).
//  * The BVH walk uses the reference slab predicate (AABB.cpp:15-35) on the
//    reference's own boxes, so a leaf is reachable exactly when the reference
//    would reach it; on top of that it orders children near-first and skips a
//    child whose entry distance exceeds the current nearest hit by more than a
//    0.1% margin.  The reference visits leaves in ascending primitive order
//    and keeps a hit only when t < length_, so its answer is the
//    lexicographic minimum of (t, category, index); the GPU keeps exactly
//    that key, so the answer does not depend on the visiting order.
//  * The grid walk and the naive scan visit in the reference order and use
//    the reference's strict t < length_ rule.
#pragma once
#include "mrt_cwt.hpp"
#include "mrt_scene.hpp"

namespace mrt {

struct RayT {
    f3 o, d, inv;
    uint32_t excl;
};

struct HitT {
    float t;
    uint32_t id;
    float u, v;
};

MRT_HD RayT make_ray(f3 o, f3 d, uint32_t excl) {
    RayT r;
    r.o = o;
    r.d = d;
    r.inv = f3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    r.excl = excl;
    return r;
}

// Traversal statistics (instrumented launches only), per query kind:
// 0 closest hit over planes/spheres/triangles, 1 lights, 2 any hit.
struct TraceStats {
    uint32_t pairs[3];
    uint32_t prims[3];
    uint32_t lane_iters;  // persistent trace: iterations a lane held a ray
    uint32_t lane_slots;  // persistent trace: iterations x lanes
};

// A word of LDS; the explicit address space keeps stack accesses on ds_*
// instructions instead of flat ones once the pointer sits in a struct.
typedef __attribute__((address_space(3))) uint32_t LdsWord;

// Per-lane traversal stack: the first kLds entries live in LDS (stride =
// block size, so lanes hit distinct banks), deeper entries in a global spill
// area sized on the host from the tree depth.
struct Stack {
    LdsWord* lds;
    uint32_t* spill;
    uint32_t lds_stride;
    uint32_t spill_stride;
    uint32_t cap_lds;
    uint32_t cap_total;
    uint32_t sp;
    uint32_t overflow;

    __device__ __forceinline__ void push(uint32_t v) {
        if (sp < cap_lds) {
            lds[sp * lds_stride] = v;
        } else if (sp < cap_total) {
            spill[(sp - cap_lds) * spill_stride] = v;
        } else {
            overflow = 1u;
            return;
        }
        ++sp;
    }
    __device__ __forceinline__ uint32_t pop() {
        --sp;
        return sp < cap_lds ? lds[sp * lds_stride] : spill[(sp - cap_lds) * spill_stride];
    }
};

MRT_HD bool key_better(float t, uint32_t id, const HitT& h) {
    return (t < h.t) || (t == h.t && id < h.id && h.id != MRT_ID_NONE);
}

// ------------------------------------------------------------- primitives --
struct TriG {
    f3 a, ab, ac;
};
__device__ __forceinline__ TriG load_tri(const float4* geo, uint32_t i) {
    const float4 g0 = geo[3u * i + 0u];
    const float4 g1 = geo[3u * i + 1u];
    const float4 g2 = geo[3u * i + 2u];
    TriG t;
    t.a = f3{g0.x, g0.y, g0.z};
    t.ab = f3{g0.w, g1.x, g1.y};
    t.ac = f3{g1.z, g1.w, g2.x};
    return t;
}

// Triangle::intersect(Intersection*, Ray) without the length_ test: returns
// true and (t, u, v) when the ray meets the triangle at t >= Epsilon.
MRT_HD bool tri_test(const TriG& g, const RayT& r, float* tt, float* uu, float* vv) {
    const f3 perp = cross(r.d, g.ac);
    const float det = dot(g.ab, perp);
    if (fabsf(det) < kEpsilon) return false;
    const float inv = 1.0f / det;
    const f3 s = sub(r.o, g.a);
    const float u = inv * dot(s, perp);
    if (u < 0.0f || u > 1.0f) return false;
    const f3 q = cross(s, g.ab);
    const float v = inv * dot(r.d, q);
    if (v < 0.0f || (u + v) > 1.0f) return false;
    const float t = inv * dot(g.ac, q);
    *tt = t;
    *uu = u;
    *vv = v;
    return !(t < kEpsilon);
}
// Triangle::intersect(Ray, dist), Triangle.cpp:105-136.
MRT_HD bool tri_any(const TriG& g, const RayT& r, float dist) {
    const f3 perp = cross(r.d, g.ac);
    const float det = dot(g.ab, perp);
    if (fabsf(det) < kEpsilon) return false;
    const float inv = 1.0f / det;
    const f3 s = sub(r.o, g.a);
    const float u = inv * dot(s, perp);
    if (u < 0.0f || u > 1.0f) return false;
    const f3 q = cross(s, g.ab);
    const float v = inv * dot(r.d, q);
    if (v < 0.0f || (u + v) > 1.0f) return false;
    const float t = inv * dot(g.ac, q);
    return t > kEpsilon && t < dist;
}

// Sphere::intersect root, Sphere.cpp:18-41.
MRT_HD bool sph_root(f3 c, float sqr, const RayT& r, float* tt) {
    const f3 oc = sub(c, r.o);
    const float proj = dot(oc, r.d);
    const float ocm = length(oc);
    const float A = dot(r.d, r.d);
    const float B = 2.0f * -proj;
    const float C = ocm * ocm - sqr;
    const float disc = B * B - (4.0f * A) * C;
    if (disc < 0.0f) return false;
    const float rd = sqrtf(disc);
    const float t1 = -B + rd;
    const float t2 = -B - rd;
    *tt = smin(t1, t2) / (2.0f * A);
    return true;
}

// Plane::intersect distance, Plane.cpp:17-45.
MRT_HD bool pla_t(f3 n, f3 p, const RayT& r, float* tt) {
    const float np = dot(n, r.d);
    if (fabsf(np) < kEpsilon) return false;
    const f3 vp = sub(p, r.o);
    const float sp = dot(n, vp);
    *tt = sp / np;
    return true;
}

// Closest-hit test of triangle `idx` whose data g is already loaded, under
// the reference's visit-order rule (t < best), as prim_closest<TRIANGLE, true>.
__device__ __forceinline__ bool tri_closest_strict(const TriG& g, uint32_t idx, const RayT& r, HitT& h) {
    const uint32_t id = MRT_MAKE_ID(MRT_CAT_TRIANGLE, idx);
    float t, u = 0.0f, v = 0.0f;
    if (id == r.excl || !tri_test(g, r, &t, &u, &v) || !(t < h.t)) return false;
    h.t = t;
    h.id = id;
    h.u = u;
    h.v = v;
    return true;
}

// Closest-hit test of primitive `idx` of category CAT.  STRICT selects the
// reference's visit-order rule (t < best); otherwise the (t, id) key.
template <int CAT, bool STRICT>
__device__ __forceinline__ bool prim_closest(const SceneDev& s, uint32_t idx, const RayT& r, HitT& h) {
    const uint32_t id = MRT_MAKE_ID(CAT, idx);
    float t, u = 0.0f, v = 0.0f;
    bool ok;
    if (CAT == MRT_CAT_TRIANGLE) {
        if (id == r.excl) return false;
        ok = tri_test(load_tri(s.tri_geo, idx), r, &t, &u, &v);
    } else if (CAT == MRT_CAT_LIGHT) {
        if (id == r.excl) return false;
        ok = tri_test(load_tri(s.light_geo, idx), r, &t, &u, &v);
    } else if (CAT == MRT_CAT_SPHERE) {
        const float4 c = s.sph[idx];
        ok = sph_root(f3{c.x, c.y, c.z}, c.w, r, &t) && !(t < 1.0e-05f);
    } else {
        if (id == r.excl) return false;
        const float4 n = s.pla[2u * idx];
        const float4 p = s.pla[2u * idx + 1u];
        ok = pla_t(f3{n.x, n.y, n.z}, f3{p.x, p.y, p.z}, r, &t) && !(t < kEpsilon);
    }
    if (!ok) return false;
    const bool better = STRICT ? (t < h.t) : key_better(t, id, h);
    if (better) {
        h.t = t;
        h.id = id;
        h.u = u;
        h.v = v;
    }
    return better;
}

template <int CAT>
__device__ __forceinline__ bool prim_any(const SceneDev& s, uint32_t idx, const RayT& r, float dist) {
    const uint32_t id = MRT_MAKE_ID(CAT, idx);
    if (CAT == MRT_CAT_TRIANGLE) {
        if (id == r.excl) return false;
        return tri_any(load_tri(s.tri_geo, idx), r, dist);
    } else if (CAT == MRT_CAT_SPHERE) {
        const float4 c = s.sph[idx];
        float t;
        if (!sph_root(f3{c.x, c.y, c.z}, c.w, r, &t)) return false;
        return t > 1.0e-05f && t < dist;
    } else {
        if (id == r.excl) return false;
        const float4 n = s.pla[2u * idx];
        const float4 p = s.pla[2u * idx + 1u];
        float t;
        if (!pla_t(f3{n.x, n.y, n.z}, f3{p.x, p.y, p.z}, r, &t)) return false;
        return t > kEpsilon && t < dist;
    }
}

// ------------------------------------------------------------------- BVH --
MRT_HD float cull_limit(float t) { return t * 1.001f + 1.0e-5f; }

template <int CAT, bool COUNT>
__device__ void bvh_closest(const SceneDev& s, const Accel& a, const RayT& r, HitT& h, Stack& st, TraceStats& ts) {
    float e;
    if (!slab(f3{a.root_mn[0], a.root_mn[1], a.root_mn[2]}, f3{a.root_mx[0], a.root_mx[1], a.root_mx[2]}, r.o, r.inv,
              &e))
        return;
    if (a.root_count > 0u) {
        for (uint32_t k = 0; k < a.root_count; ++k) {
            const uint32_t slot = a.root_ref + k;
            if (COUNT) ++ts.prims[CAT == MRT_CAT_LIGHT ? 1 : 0];
            prim_closest<CAT, false>(s, a.perm ? a.perm[slot] : slot, r, h);
        }
        return;
    }
    st.sp = 0;
    uint32_t node = 0;
    while (true) {
        const BNode* nd = a.nodes + node;
        const float4 q0 = nd->q0;
        const float4 q1 = nd->q1;
        const float4 q2 = nd->q2;
        const uint4 q3 = nd->q3;
        if (COUNT) ++ts.pairs[CAT == MRT_CAT_LIGHT ? 1 : 0];
        float eL, eR;
        const float lim = cull_limit(h.t);
        bool hL = slab(f3{q0.x, q0.y, q0.z}, f3{q0.w, q1.x, q1.y}, r.o, r.inv, &eL) && eL <= lim;
        bool hR = slab(f3{q1.z, q1.w, q2.x}, f3{q2.y, q2.z, q2.w}, r.o, r.inv, &eR) && eR <= lim;
        if (hL && q3.y > 0u) {
            for (uint32_t k = 0; k < q3.y; ++k) {
                const uint32_t slot = q3.x + k;
                if (COUNT) ++ts.prims[CAT == MRT_CAT_LIGHT ? 1 : 0];
                prim_closest<CAT, false>(s, a.perm ? a.perm[slot] : slot, r, h);
            }
            hL = false;
        }
        if (hR && q3.w > 0u) {
            for (uint32_t k = 0; k < q3.w; ++k) {
                const uint32_t slot = q3.z + k;
                if (COUNT) ++ts.prims[CAT == MRT_CAT_LIGHT ? 1 : 0];
                prim_closest<CAT, false>(s, a.perm ? a.perm[slot] : slot, r, h);
            }
            hR = false;
        }
        if (hL && hR) {
            const bool leftFirst = eL <= eR;
            node = leftFirst ? q3.x : q3.z;
            st.push(leftFirst ? q3.z : q3.x);
        } else if (hL) {
            node = q3.x;
        } else if (hR) {
            node = q3.z;
        } else {
            if (st.sp == 0u) break;
            node = st.pop();
        }
    }
}

template <int CAT, bool COUNT>
__device__ bool bvh_any(const SceneDev& s, const Accel& a, const RayT& r, float dist, Stack& st, TraceStats& ts) {
    float e;
    if (!slab(f3{a.root_mn[0], a.root_mn[1], a.root_mn[2]}, f3{a.root_mx[0], a.root_mx[1], a.root_mx[2]}, r.o, r.inv,
              &e))
        return false;
    if (a.root_count > 0u) {
        for (uint32_t k = 0; k < a.root_count; ++k) {
            if (COUNT) ++ts.prims[2];
            if (prim_any<CAT>(s, a.root_ref + k, r, dist)) return true;
        }
        return false;
    }
    const float lim = cull_limit(dist);
    st.sp = 0;
    uint32_t node = 0;
    while (true) {
        const BNode* nd = a.nodes + node;
        const float4 q0 = nd->q0;
        const float4 q1 = nd->q1;
        const float4 q2 = nd->q2;
        const uint4 q3 = nd->q3;
        if (COUNT) ++ts.pairs[2];
        float eL, eR;
        bool hL = slab(f3{q0.x, q0.y, q0.z}, f3{q0.w, q1.x, q1.y}, r.o, r.inv, &eL) && eL <= lim;
        bool hR = slab(f3{q1.z, q1.w, q2.x}, f3{q2.y, q2.z, q2.w}, r.o, r.inv, &eR) && eR <= lim;
        if (hL && q3.y > 0u) {
            for (uint32_t k = 0; k < q3.y; ++k) {
                if (COUNT) ++ts.prims[2];
                if (prim_any<CAT>(s, q3.x + k, r, dist)) return true;
            }
            hL = false;
        }
        if (hR && q3.w > 0u) {
            for (uint32_t k = 0; k < q3.w; ++k) {
                if (COUNT) ++ts.prims[2];
                if (prim_any<CAT>(s, q3.z + k, r, dist)) return true;
            }
            hR = false;
        }
        if (hL && hR) {
            const bool leftFirst = eL <= eR;
            node = leftFirst ? q3.x : q3.z;
            st.push(leftFirst ? q3.z : q3.x);
        } else if (hL) {
            node = q3.x;
        } else if (hR) {
            node = q3.z;
        } else {
            if (st.sp == 0u) break;
            node = st.pop();
        }
    }
    return false;
}

// ------------------------------------------------------------ RegularGrid --
struct Dda {
    int32_t X, Y, Z, sx, sy, sz, ox, oy, oz;
    float tx, ty, tz, dx, dy, dz;
};

// 3D-DDA setup, RegularGrid.hpp:161-231.
__device__ __forceinline__ void dda_setup(const Accel& a, const RayT& r, Dda& g) {
    const f3 mn = f3{a.ext_mn[0], a.ext_mn[1], a.ext_mn[2]};
    const f3 c = mul(sub(r.o, mn), f3{a.sr[0], a.sr[1], a.sr[2]});
    const int32_t gm = a.grid_size - 1;
    g.X = min(gm, max(0, static_cast<int32_t>(c.x)));
    g.Y = min(gm, max(0, static_cast<int32_t>(c.y)));
    g.Z = min(gm, max(0, static_cast<int32_t>(c.z)));
    float cbx, cby, cbz;
    if (r.d.x > 0.0f) { g.sx = 1; g.ox = a.grid_size; cbx = mn.x + static_cast<float>(g.X + 1) * a.cw[0]; }
    else { g.sx = -1; g.ox = -1; cbx = mn.x + static_cast<float>(g.X) * a.cw[0]; }
    if (r.d.y > 0.0f) { g.sy = 1; g.oy = a.grid_size; cby = mn.y + static_cast<float>(g.Y + 1) * a.cw[1]; }
    else { g.sy = -1; g.oy = -1; cby = mn.y + static_cast<float>(g.Y) * a.cw[1]; }
    if (r.d.z > 0.0f) { g.sz = 1; g.oz = a.grid_size; cbz = mn.z + static_cast<float>(g.Z + 1) * a.cw[2]; }
    else { g.sz = -1; g.oz = -1; cbz = mn.z + static_cast<float>(g.Z) * a.cw[2]; }
    g.dx = g.dy = g.dz = 0.0f;
    if (fabsf(r.d.x) > kEpsilon) { const float rr = 1.0f / r.d.x; g.tx = (cbx - r.o.x) * rr; g.dx = (a.cw[0] * static_cast<float>(g.sx)) * rr; }
    else g.tx = kRayLengthMax;
    if (fabsf(r.d.y) > kEpsilon) { const float rr = 1.0f / r.d.y; g.ty = (cby - r.o.y) * rr; g.dy = (a.cw[1] * static_cast<float>(g.sy)) * rr; }
    else g.ty = kRayLengthMax;
    if (fabsf(r.d.z) > kEpsilon) { const float rr = 1.0f / r.d.z; g.tz = (cbz - r.o.z) * rr; g.dz = (a.cw[2] * static_cast<float>(g.sz)) * rr; }
    else g.tz = kRayLengthMax;
}
__device__ __forceinline__ uint32_t dda_cell(const Accel& a, const Dda& g) {
    return static_cast<uint32_t>(g.X) + (static_cast<uint32_t>(g.Y) << a.grid_shift) +
           (static_cast<uint32_t>(g.Z) << (a.grid_shift * 2u));
}
// One DDA step; with `limit` the phase-2 test `length_ < tmax` comes first.
// Returns false when the walk ends.
__device__ __forceinline__ bool dda_step(Dda& g, bool useLimit, float limit) {
    if (g.tx < g.ty) {
        if (g.tx < g.tz) {
            if (useLimit && limit < g.tx) return false;
            g.X += g.sx;
            if (g.X == g.ox) return false;
            g.tx = g.tx + g.dx;
        } else {
            if (useLimit && limit < g.tz) return false;
            g.Z += g.sz;
            if (g.Z == g.oz) return false;
            g.tz = g.tz + g.dz;
        }
    } else {
        if (g.ty < g.tz) {
            if (useLimit && limit < g.ty) return false;
            g.Y += g.sy;
            if (g.Y == g.oy) return false;
            g.ty = g.ty + g.dy;
        } else {
            if (useLimit && limit < g.tz) return false;
            g.Z += g.sz;
            if (g.Z == g.oz) return false;
            g.tz = g.tz + g.dz;
        }
    }
    return true;
}

// RegularGrid<T>::trace, RegularGrid.hpp:157-337.
template <int CAT, bool COUNT>
__device__ void grid_closest(const SceneDev& s, const Accel& a, const RayT& r, HitT& h, TraceStats& ts) {
    Dda g;
    dda_setup(a, r, g);
    bool phase2 = false;
    while (true) {
        const uint32_t c = dda_cell(a, g);
        const uint32_t b = a.cell_off[c], e = a.cell_off[c + 1u];
        if (!phase2) {
            bool hit = false;
            for (uint32_t k = b; k < e; ++k) {
                if (COUNT) ++ts.prims[0];
                if (prim_closest<CAT, true>(s, a.cell_ref[k], r, h)) { hit = true; break; }
            }
            if (hit) { phase2 = true; continue; }
            if (COUNT) ++ts.pairs[0];
            if (!dda_step(g, false, 0.0f)) return;
        } else {
            for (uint32_t k = b; k < e; ++k) {
                if (COUNT) ++ts.prims[0];
                prim_closest<CAT, true>(s, a.cell_ref[k], r, h);
            }
            if (COUNT) ++ts.pairs[0];
            if (!dda_step(g, true, h.t)) return;
        }
    }
}

// RegularGrid<T>::shadowTrace, RegularGrid.hpp:340-465.
template <int CAT, bool COUNT>
__device__ bool grid_any(const SceneDev& s, const Accel& a, const RayT& r, float dist, TraceStats& ts) {
    Dda g;
    dda_setup(a, r, g);
    while (true) {
        const uint32_t c = dda_cell(a, g);
        const uint32_t b = a.cell_off[c], e = a.cell_off[c + 1u];
        for (uint32_t k = b; k < e; ++k) {
            if (COUNT) ++ts.prims[2];
            if (prim_any<CAT>(s, a.cell_ref[k], r, dist)) return true;
        }
        if (COUNT) ++ts.pairs[2];
        if (!dda_step(g, false, 0.0f)) return false;
    }
}

// ------------------------------------------------------------ dispatchers --
template <int CAT, bool COUNT>
__device__ __forceinline__ void accel_closest(const SceneDev& s, const Accel& a, const RayT& r, HitT& h, Stack& st,
                                              TraceStats& ts) {
    if (a.count == 0u) return;
    if (a.kind == MRT_BVH) {
        bvh_closest<CAT, COUNT>(s, a, r, h, st, ts);
    } else if (a.kind == MRT_REGULAR_GRID) {
        grid_closest<CAT, COUNT>(s, a, r, h, ts);
    } else {
        for (uint32_t i = 0; i < a.count; ++i) {
            if (COUNT) ++ts.prims[0];
            prim_closest<CAT, true>(s, i, r, h);
        }
    }
}
template <int CAT, bool COUNT>
__device__ __forceinline__ bool accel_any(const SceneDev& s, const Accel& a, const RayT& r, float dist, Stack& st,
                                          TraceStats& ts) {
    if (a.count == 0u) return false;
    if (a.kind == MRT_BVH) return bvh_any<CAT, COUNT>(s, a, r, dist, st, ts);
    if (a.kind == MRT_REGULAR_GRID) return grid_any<CAT, COUNT>(s, a, r, dist, ts);
    for (uint32_t i = 0; i < a.count; ++i) {
        if (COUNT) ++ts.prims[2];
        if (prim_any<CAT>(s, i, r, dist)) return true;
    }
    return false;
}

// Scene::traceLights (Scene.cpp:31-37): every area light in index order, or
// the conservative light BVH with the (t, id) key when there are many.
template <bool COUNT>
__device__ __forceinline__ void lights_closest(const SceneDev& s, const RayT& r, HitT& h, Stack& st, TraceStats& ts) {
    if (s.n_area_lights == 0u) return;
    if (s.lights_linear) {
        for (uint32_t i = 0; i < s.n_lights; ++i) {
            if (s.light_info[i].kind != MRT_LIGHT_AREA) continue;
            if (COUNT) ++ts.prims[1];
            prim_closest<MRT_CAT_LIGHT, true>(s, i, r, h);
        }
    } else {
        bvh_closest<MRT_CAT_LIGHT, COUNT>(s, s.light_acc, r, h, st, ts);
    }
}

// Full closest-hit query.  GENERIC = false compiles only the triangle and
// light paths (scenes without planes or spheres).
template <bool GENERIC, bool COUNT>
__device__ __forceinline__ HitT closest_hit(const SceneDev& s, const RayT& r, float tmax, Stack& st, TraceStats& ts) {
    HitT h;
    h.t = tmax;
    h.id = MRT_ID_NONE;
    h.u = 0.0f;
    h.v = 0.0f;
    if (GENERIC) {
        accel_closest<MRT_CAT_PLANE, COUNT>(s, s.acc[MRT_CAT_PLANE], r, h, st, ts);
        accel_closest<MRT_CAT_SPHERE, COUNT>(s, s.acc[MRT_CAT_SPHERE], r, h, st, ts);
    }
    accel_closest<MRT_CAT_TRIANGLE, COUNT>(s, s.acc[MRT_CAT_TRIANGLE], r, h, st, ts);
    lights_closest<COUNT>(s, r, h, st, ts);
    return h;
}

// ----------------------------------------------------------- combined walk --
// Persistent kernels walk one tree that joins the triangle BVH and the light
// BVH under a virtual root pair (layout in mrt_scene.hpp), one unit of work
// per lane and call: a node visit and / or one leaf record's test (its exact
// box and up to two triangles).  Leaves a visit reaches are queued (at most
// four per lane) and tested later,
// so a wave runs visits and tests as separate, mostly convergent phases and
// a lane can take a new ray the moment its ray is done.  The slab predicate on
// the reference's own boxes (the virtual root tests the triangle root box as
// BVH<T>::trace does, BVH.hpp:164-166), the culling margin and the (t, id)
// key are those of bvh_closest / bvh_any, so the answer does not depend on
// the visiting order.  Area lights of a scene with at most 16 of them sit in
// one leaf under an infinite box: Scene::traceLights tests every one of them
// (Scene.cpp:31-37), and under the (t, id) key a scan in any order equals the
// reference's ascending scan with t < length_.
constexpr uint32_t kNoNode = 0xFFFFFFFFu;

// Per-lane walk state.  Leaf words are (count << 27 | slot), 0 = none: `cur`
// is the leaf whose triangles are being tested, l0..l3 the leaves the last
// visit reached (by child slot, zeros allowed).  A lane visits again only
// once l0..l3 are empty, and `cur` is refilled from them as it runs out, so
// cur == 0 means no triangle is pending.
struct Walk {
    RayT r;
    HitT h;
    uint32_t node;  // next node to visit, kNoNode when none is left
    uint32_t cur, l0, l1, l2, l3;
};

__device__ __forceinline__ bool inv_finite(const RayT& r) {
    return fabsf(r.inv.x) < INFINITY && fabsf(r.inv.y) < INFINITY && fabsf(r.inv.z) < INFINITY;
}

__device__ __forceinline__ void walk_clear(Walk& w) {
    w.node = kNoNode;
    w.cur = w.l0 = w.l1 = w.l2 = w.l3 = 0u;
}

__device__ __forceinline__ void walk_begin(Walk& w, uint32_t root, float tmax, Stack& st) {
    w.h.t = tmax;
    w.h.id = MRT_ID_NONE;
    w.h.u = 0.0f;
    w.h.v = 0.0f;
    walk_clear(w);
    w.node = root;
    st.sp = 0;
}

__device__ __forceinline__ bool walk_done(const Walk& w) { return w.node == kNoNode && w.cur == 0u; }
__device__ __forceinline__ bool walk_can_visit(const Walk& w) {
    return w.node != kNoNode && (w.l0 | w.l1 | w.l2 | w.l3) == 0u;
}

// cur <- the first reached leaf when cur is empty (branch free).
__device__ __forceinline__ void walk_pull(Walk& w) {
    const bool need = w.cur == 0u;
    const bool t0 = need && w.l0 != 0u;
    const bool t1 = need && !t0 && w.l1 != 0u;
    const bool t2 = need && !t0 && !t1 && w.l2 != 0u;
    const bool t3 = need && !t0 && !t1 && !t2 && w.l3 != 0u;
    w.cur = t0 ? w.l0 : (t1 ? w.l1 : (t2 ? w.l2 : (t3 ? w.l3 : w.cur)));
    w.l0 = t0 ? 0u : w.l0;
    w.l1 = t1 ? 0u : w.l1;
    w.l2 = t2 ? 0u : w.l2;
    w.l3 = t3 ? 0u : w.l3;
}

// Slab test for rays whose inverse direction is finite in every axis: then no
// (bound - origin) * inv product is NaN (boxes are finite, or infinite with a
// finite origin, or all-NaN), so std::min / std::max agree with IEEE min /
// max and this returns what slab() returns, with fewer instructions.  An
// all-NaN box yields tmax = NaN and is rejected by both.
MRT_HD bool slab_finite(const f3& mn, const f3& mx, const f3& o, const f3& inv, float* entry) {
    const float x1 = (mn.x - o.x) * inv.x, x2 = (mx.x - o.x) * inv.x;
    const float y1 = (mn.y - o.y) * inv.y, y2 = (mx.y - o.y) * inv.y;
    const float z1 = (mn.z - o.z) * inv.z, z2 = (mx.z - o.z) * inv.z;
    const float tmin = fmaxf(fmaxf(fminf(x1, x2), fminf(y1, y2)), fmaxf(fminf(z1, z2), 0.0f));
    const float tmax = fminf(fminf(fmaxf(x1, x2), fmaxf(y1, y2)), fmaxf(z1, z2));
    *entry = tmin;
    return tmax >= tmin;
}

// One pair visit: queues reached leaves, moves to the nearer reached interior
// child (pushing the other) or pops the stack.
template <bool FINITE, bool COUNT>
__device__ __forceinline__ void walk_visit(const SceneDev& s, Walk& w, float lim, Stack& st, TraceStats& ts,
                                           int kind) {
    const BNode* nd = s.walk_nodes + w.node;
    const float4 q0 = nd->q0;
    const float4 q1 = nd->q1;
    const float4 q2 = nd->q2;
    const uint4 q3 = nd->q3;
    if (COUNT && w.node >= kWalkFirstPair) ++ts.pairs[kind == 2 ? 2 : (w.node >= s.walk_light_pair ? 1 : 0)];
    float eL, eR;
    bool hL, hR;
    if (FINITE) {
        hL = slab_finite(f3{q0.x, q0.y, q0.z}, f3{q0.w, q1.x, q1.y}, w.r.o, w.r.inv, &eL);
        hR = slab_finite(f3{q1.z, q1.w, q2.x}, f3{q2.y, q2.z, q2.w}, w.r.o, w.r.inv, &eR);
    } else {
        hL = slab(f3{q0.x, q0.y, q0.z}, f3{q0.w, q1.x, q1.y}, w.r.o, w.r.inv, &eL);
        hR = slab(f3{q1.z, q1.w, q2.x}, f3{q2.y, q2.z, q2.w}, w.r.o, w.r.inv, &eR);
    }
    hL = hL && eL <= lim;
    hR = hR && eR <= lim;
    const bool leafL = q3.y != 0u, leafR = q3.w != 0u;
    w.l0 = (hL && leafL) ? q3.y : 0u;
    w.l1 = (hR && leafR) ? q3.w : 0u;
    const bool inL = hL && !leafL, inR = hR && !leafR;
    const bool leftFirst = eL <= eR;
    uint32_t next = (inL && (!inR || leftFirst)) ? q3.x : q3.z;
    if (inL && inR) {
        st.push(leftFirst ? q3.z : q3.x);
    } else if (!inL && !inR) {
        next = st.sp == 0u ? kNoNode : st.pop();
    }
    w.node = next;
    walk_pull(w);
}

// Orders (key, ref) pairs so that ka <= kb.
__device__ __forceinline__ void cswap(float& ka, uint32_t& ra, float& kb, uint32_t& rb) {
    const bool sw = kb < ka;
    const float k = sw ? kb : ka;
    const uint32_t r = sw ? rb : ra;
    kb = sw ? ka : kb;
    rb = sw ? ra : rb;
    ka = k;
    ra = r;
}

// One compressed-node visit (layout in mrt_cwt.hpp), the walk_visit_wide of
// the quantized boxes: open slots are entered at distance 0, reached leaves
// go to l0..l3 (their exact boxes are tested with their triangles), reached
// interior slots are ordered by entry distance.
struct CwNode {
    uint4 n0, n1, n2, n3;
};
__device__ __forceinline__ CwNode load_cw(const SceneDev& s, uint32_t node) {
    const uint4* p = s.cw_nodes + 4u * node;
    return CwNode{p[0], p[1], p[2], p[3]};
}

__device__ __forceinline__ bool cw_slot(const CwNode& c, uint32_t k, const f3& org, const f3& sc, const RayT& r,
                                        float lim, float* e) {
    const f3 mn{cw_deq(org.x, cw_byte(c.n1.x, k), sc.x), cw_deq(org.y, cw_byte(c.n1.y, k), sc.y),
                cw_deq(org.z, cw_byte(c.n1.z, k), sc.z)};
    const f3 mx{cw_deq(org.x, cw_byte(c.n1.w, k), sc.x), cw_deq(org.y, cw_byte(c.n2.x, k), sc.y),
                cw_deq(org.z, cw_byte(c.n2.y, k), sc.z)};
    const uint32_t kind = (c.n0.w >> (24u + 2u * k)) & 3u;
    const bool open = (c.n3.z >> k) & 1u;
    const bool hit = slab_finite(mn, mx, r.o, r.inv, e) && *e <= lim;
    *e = open ? 0.0f : *e;
    return kind != kCwEmpty && (open || hit);
}

template <bool COUNT>
__device__ __forceinline__ void cw_visit(const SceneDev& s, Walk& w, const CwNode& c, float lim, Stack& st,
                                         TraceStats& ts, int kind) {
    if (COUNT) ++ts.pairs[kind == 2 ? 2 : (w.node >= s.wide_light_node ? 1 : 0)];
    const f3 org{__uint_as_float(c.n0.x), __uint_as_float(c.n0.y), __uint_as_float(c.n0.z)};
    const f3 sc{cw_scale(c.n0.w & 0xFFu), cw_scale((c.n0.w >> 8) & 0xFFu), cw_scale((c.n0.w >> 16) & 0xFFu)};
    float e0, e1, e2, e3;
    const bool h0 = cw_slot(c, 0u, org, sc, w.r, lim, &e0);
    const bool h1 = cw_slot(c, 1u, org, sc, w.r, lim, &e1);
    const bool h2 = cw_slot(c, 2u, org, sc, w.r, lim, &e2);
    const bool h3 = cw_slot(c, 3u, org, sc, w.r, lim, &e3);
    const uint32_t kinds = c.n0.w >> 24;
    const bool f0 = ((kinds >> 0) & 3u) == kCwLeaf, f1 = ((kinds >> 2) & 3u) == kCwLeaf;
    const bool f2 = ((kinds >> 4) & 3u) == kCwLeaf, f3v = ((kinds >> 6) & 3u) == kCwLeaf;
    uint32_t n0 = c.n2.z, n1 = c.n2.w, n2 = c.n3.x, n3 = c.n3.y;
    w.l0 = (h0 && f0) ? n0 : 0u;
    w.l1 = (h1 && f1) ? n1 : 0u;
    w.l2 = (h2 && f2) ? n2 : 0u;
    w.l3 = (h3 && f3v) ? n3 : 0u;
    float k0 = (h0 && !f0) ? e0 : INFINITY, k1 = (h1 && !f1) ? e1 : INFINITY;
    float k2 = (h2 && !f2) ? e2 : INFINITY, k3 = (h3 && !f3v) ? e3 : INFINITY;
    cswap(k0, n0, k1, n1);
    cswap(k2, n2, k3, n3);
    cswap(k0, n0, k2, n2);
    cswap(k1, n1, k3, n3);
    cswap(k1, n1, k2, n2);
    if (k3 < INFINITY) st.push(n3);
    if (k2 < INFINITY) st.push(n2);
    if (k1 < INFINITY) st.push(n1);
    if (k0 < INFINITY) w.node = n0;
    else w.node = st.sp == 0u ? kNoNode : st.pop();
    walk_pull(w);
}


// tri_test without early exits (the same operations; lanes of a wave rarely
// agree on an exit, so the branches would only cost mask bookkeeping).
__device__ __forceinline__ bool tri_test_flat(const TriG& g, const RayT& r, float* tt, float* uu, float* vv) {
    const f3 perp = cross(r.d, g.ac);
    const float det = dot(g.ab, perp);
    const float inv = 1.0f / det;
    const f3 s = sub(r.o, g.a);
    const float u = inv * dot(s, perp);
    const f3 q = cross(s, g.ab);
    const float v = inv * dot(r.d, q);
    const float t = inv * dot(g.ac, q);
    *tt = t;
    *uu = u;
    *vv = v;
    return !(fabsf(det) < kEpsilon) && !(u < 0.0f || u > 1.0f) && !(v < 0.0f || (u + v) > 1.0f) &&
           !(t < kEpsilon);
}

template <bool GENERIC, bool COUNT>
__device__ __forceinline__ bool any_hit(const SceneDev& s, const RayT& r, float dist, Stack& st, TraceStats& ts) {
    if (GENERIC) {
        if (accel_any<MRT_CAT_PLANE, COUNT>(s, s.acc[MRT_CAT_PLANE], r, dist, st, ts)) return true;
        if (accel_any<MRT_CAT_SPHERE, COUNT>(s, s.acc[MRT_CAT_SPHERE], r, dist, st, ts)) return true;
    }
    return accel_any<MRT_CAT_TRIANGLE, COUNT>(s, s.acc[MRT_CAT_TRIANGLE], r, dist, st, ts);
}

}  // namespace mrt
