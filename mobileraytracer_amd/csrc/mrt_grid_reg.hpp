// RegularGrid cell registration shared by the host and GPU grid builds: the
// candidate cell range of a primitive box (RegularGrid.hpp:94-132) and the
// primitives' intersectBox tests (Triangle.cpp:154-246, Sphere.cpp:102-123,
// Plane.cpp:106-155), statement for statement, so that both builds register
// the same cells.
#pragma once
#include "mrt_math.hpp"

namespace mrt {

// getAABB of the three primitive kinds (Triangle.cpp:145-152,
// Sphere.cpp:91-96, Plane.cpp:80-104).
MRT_HD Box triangle_box_of(f3 A, f3 AB, f3 AC) {
    const f3 B = add(A, AB);
    const f3 C = add(A, AC);
    return Box{vmin(A, vmin(B, C)), vmax(A, vmax(B, C))};
}

MRT_HD Box sphere_box_of(f3 c, float sq_radius) {
    const float r = sqrtf(sq_radius);
    return Box{subs(c, r), adds(c, r)};
}

MRT_HD Box plane_box_of(f3 n, f3 pt) {
    f3 right{0.0f, 0.0f, 0.0f};
    if (n.x >= 1.0f) right = f3{0, 1, 1};
    else if (n.y >= 1.0f) right = f3{1, 0, 1};
    else if (n.z >= 1.0f) right = f3{1, 1, 0};
    else if (n.x <= -1.0f) right = f3{0, 1, 1};
    else if (n.y <= -1.0f) right = f3{1, 0, 1};
    else if (n.z <= -1.0f) right = f3{1, 1, 0};
    right = normalize(right);
    return Box{add(pt, muls(right, -100.0f)), add(pt, muls(right, 100.0f))};
}

// Cell layout of a grid: gs cells per axis over bounds.
struct GridGeom {
    Box bounds;
    int32_t gs;
    float size[3];  // bounds extent per axis
    float d[3];     // cell size per axis
    float rec[3];   // 1 / cell size (1 for a flat axis)
};

MRT_HD GridGeom grid_geom(const Box& bounds, int32_t gs) {
    GridGeom g;
    g.bounds = bounds;
    g.gs = gs;
    const float n = static_cast<float>(gs);
    g.size[0] = bounds.mx.x - bounds.mn.x;
    g.size[1] = bounds.mx.y - bounds.mn.y;
    g.size[2] = bounds.mx.z - bounds.mn.z;
    for (int a = 0; a < 3; ++a) {
        g.d[a] = g.size[a] / n;
        g.rec[a] = g.d[a] > 0.0f ? 1.0f / g.d[a] : 1.0f;
    }
    return g;
}

// Candidate cells [lo, hi] of a box on one axis.
MRT_HD void grid_range(const GridGeom& g, int a, float lo, float hi, int32_t* c1, int32_t* c2) {
    const float mn = a == 0 ? g.bounds.mn.x : a == 1 ? g.bounds.mn.y : g.bounds.mn.z;
    int32_t x1 = static_cast<int32_t>((lo - mn) * g.rec[a]);
    int32_t x2 = static_cast<int32_t>((hi - mn) * g.rec[a]) + 1;
    x1 = x1 > 0 ? x1 : 0;
    x2 = x2 < g.gs - 1 ? x2 : g.gs - 1;
    x2 = fabsf(g.size[a]) < kEpsilon ? 0 : x2;
    x1 = x1 < x2 ? x1 : x2;
    *c1 = x1;
    *c2 = x2;
}

MRT_HD Box grid_cell(const GridGeom& g, int32_t x, int32_t y, int32_t z) {
    const f3 pos{g.bounds.mn.x + static_cast<float>(x) * g.d[0], g.bounds.mn.y + static_cast<float>(y) * g.d[1],
                 g.bounds.mn.z + static_cast<float>(z) * g.d[2]};
    return Box{pos, add(pos, f3{g.d[0], g.d[1], g.d[2]})};
}

// Triangle::intersectBox, Triangle.cpp:154-246: an edge of the triangle
// crosses the box, the box diagonal hits the triangle, or the triangle lies
// in the diagonal's plane.
MRT_HD bool triangle_box_overlap(f3 A, f3 AB, f3 AC, const Box& box) {
    auto rayBox = [&](f3 orig, f3 vec) -> bool {
        float tNear = __FLT_MIN__;
        float tFar = __FLT_MAX__;
        for (int a = 0; a < 3; ++a) {
            const float o = a == 0 ? orig.x : a == 1 ? orig.y : orig.z;
            const float v = a == 0 ? vec.x : a == 1 ? vec.y : vec.z;
            const float bmn = a == 0 ? box.mn.x : a == 1 ? box.mn.y : box.mn.z;
            const float bmx = a == 0 ? box.mx.x : a == 1 ? box.mx.y : box.mx.z;
            if (fabsf(v) < kEpsilon) {
                if ((o < bmn) || ((o + v) > bmx)) return false;
            } else {
                float t1 = (bmn - o) / v;
                float t2 = (bmx - o) / v;
                if (t1 > t2) {
                    const float t = t1;
                    t1 = t2;
                    t2 = t;
                }
                tNear = smax(t1, tNear);
                tFar = smin(t2, tFar);
                if ((tNear > tFar) || (tFar < 0.0f)) return false;
            }
        }
        return true;
    };
    const f3 vec = sub(box.mx, box.mn);
    const bool iAB = rayBox(A, AB);
    const bool iAC = rayBox(A, AC);
    const f3 B = add(A, AB);
    const f3 C = add(A, AC);
    const bool iBC = rayBox(B, sub(C, B));
    bool iRay = false;
    {
        const f3 d = vec;
        const f3 perp = cross(d, AC);
        const float det = dot(AB, perp);
        if (!(fabsf(det) < kEpsilon)) {
            const float inv = 1.0f / det;
            const f3 s = sub(box.mn, A);
            const float u = inv * dot(s, perp);
            if (!(u < 0.0f || u > 1.0f)) {
                const f3 q = cross(s, AB);
                const float v = inv * dot(d, q);
                if (!(v < 0.0f || (u + v) > 1.0f)) {
                    const float tt = inv * dot(AC, q);
                    iRay = !(tt < kEpsilon || tt >= kRayLengthMax);
                }
            }
        }
    }
    const f3 perp = cross(vec, AC);
    const bool inside = fabsf(dot(AB, perp)) < kEpsilon;
    return iAB || iAC || iBC || iRay || inside;
}

// Sphere::intersectBox, Sphere.cpp:102-123.
MRT_HD bool sphere_box_overlap(f3 c, float sq_radius, const Box& box) {
    float dmin = 0.0f;
    for (int a = 0; a < 3; ++a) {
        const float ca = a == 0 ? c.x : a == 1 ? c.y : c.z;
        const float mn = a == 0 ? box.mn.x : a == 1 ? box.mn.y : box.mn.z;
        const float mx = a == 0 ? box.mx.x : a == 1 ? box.mx.y : box.mx.z;
        if (ca < mn) dmin = dmin + (ca - mn) * (ca - mn);
        else if (ca > mx) dmin = dmin + (ca - mx) * (ca - mx);
    }
    return dmin <= sq_radius;
}

// Plane::distance, Plane.cpp:106-121.
MRT_HD float plane_distance_to(f3 n, f3 pt, f3 q) {
    const float d = (n.x * -pt.x + n.y * -pt.y) + n.z * -pt.z;
    const float num = ((n.x * q.x + n.y * q.y) + n.z * q.z) + d;
    const float den = sqrtf((n.x * n.x + n.y * n.y) + n.z * n.z);
    return num / den;
}

// Plane::intersectBox, Plane.cpp:123-155.
MRT_HD bool plane_box_overlap(f3 n, f3 pt, const Box& box) {
    const float dP = plane_distance_to(n, pt, box.mx);
    const float dN = plane_distance_to(n, pt, box.mn);
    return (dP <= 0.0f && dN >= 0.0f) || (dP >= 0.0f && dN <= 0.0f);
}

}  // namespace mrt
