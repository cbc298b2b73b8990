// ORACLE — TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or
// calls this code; tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg use it as the checker.
//
// A CPU restatement of MobileRT's ray-scene hot path, following the reference
// sources function by function (paths relative to the MobileRT tree; every
// function below names the lines it restates).  It keeps the reference's
// algorithms and costs ("reference-faithful"): AoS triangles, the in-order SAH
// BVH walked without t-culling or near-first ordering, the linear
// Scene::traceLights, the RegularGrid walk that copies each cell's list, and
// the std::thread tile loop.  Templated on the real type R: R = float is the
// parity oracle, R = double the statistical oracle for the path tracer.
//
// glm is not available here (app/third_party is an empty submodule), so the
// vector operations are written from glm's documented generic formulas:
//   dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b)    = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
//   normalize(v)  = v * (1 / sqrt(dot(v,v)))
//   length(v)     = sqrt(dot(v,v))
//   reflect(I,N)  = I - N * dot(N,I) * 2
//   refract(I,N,e)= k = 1 - e*e*(1 - d*d); k < 0 ? 0 : e*I - (e*d + sqrt(k))*N
//   min(x,y)      = (y < x) ? y : x,   max(x,y) = (x < y) ? y : x
// The rounding of the reference's unmodified glm build is therefore not pinned
// beyond the reference's own known-answer tests; this oracle fixes the order
// above and the GPU follows the same order.
#pragma once
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "sampling.hpp"

namespace orc {

constexpr float kRayLengthMax = 1.0e+30f;  // Utils.hpp:33
constexpr float kEpsilon = 1.0e-06f;       // Utils.hpp:37
constexpr uint32_t kNumberOfBlocks = 256;  // Utils.hpp:36

// ---------------------------------------------------------------- vectors --
template <class R>
struct V3 {
    R x{}, y{}, z{};
    R& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    const R& operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> inline V3<R> v3(R x, R y, R z) { return V3<R>{x, y, z}; }
template <class R> inline V3<R> from3(const float* p) {
    return V3<R>{static_cast<R>(p[0]), static_cast<R>(p[1]), static_cast<R>(p[2])};
}
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> inline V3<R> operator-(V3<R> a, R s) { return {a.x - s, a.y - s, a.z - s}; }
template <class R> inline V3<R> operator+(V3<R> a, R s) { return {a.x + s, a.y + s, a.z + s}; }
template <class R> inline R dot(V3<R> a, V3<R> b) {
    const R xx = a.x * b.x;
    const R yy = a.y * b.y;
    const R zz = a.z * b.z;
    return (xx + yy) + zz;
}
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
template <class R> inline R length(V3<R> v) { return std::sqrt(dot(v, v)); }
template <class R> inline V3<R> normalize(V3<R> v) {
    const R inv = static_cast<R>(1) / std::sqrt(dot(v, v));
    return v * inv;
}
template <class R> inline V3<R> reflect(V3<R> i, V3<R> n) { return i - (n * dot(n, i)) * static_cast<R>(2); }
template <class R> inline V3<R> refract(V3<R> i, V3<R> n, R eta) {
    const R d = dot(n, i);
    const R k = static_cast<R>(1) - (eta * eta) * (static_cast<R>(1) - d * d);
    if (k >= static_cast<R>(0)) {
        const R f = eta * d + std::sqrt(k);
        return (eta * i) - (f * n);
    }
    return V3<R>{};
}
template <class R> inline R smin(R a, R b) { return (b < a) ? b : a; }  // std::min
template <class R> inline R smax(R a, R b) { return (a < b) ? b : a; }  // std::max
template <class R> inline V3<R> vmin(V3<R> a, V3<R> b) { return {smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)}; }
template <class R> inline V3<R> vmax(V3<R> a, V3<R> b) { return {smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)}; }
template <class R> inline bool any_gt0(V3<R> v) { return v.x > 0 || v.y > 0 || v.z > 0; }

// ------------------------------------------------------------------ AABB --
template <class R>
struct AABB {
    V3<R> mn{}, mx{};
};
// ::MobileRT::intersect(AABB, Ray), app/MobileRT/Boxes/AABB.cpp:15-35.
template <class R>
inline bool slab(const AABB<R>& box, const V3<R>& o, const V3<R>& d) {
    const R invX = static_cast<R>(1) / d.x;
    const R t1X = (box.mn.x - o.x) * invX;
    const R t2X = (box.mx.x - o.x) * invX;
    R tmin = smin(t1X, t2X);
    R tmax = smax(t1X, t2X);
    for (int axis = 1; axis < 3; ++axis) {
        const R inv = static_cast<R>(1) / d[axis];
        const R t1 = (box.mn[axis] - o[axis]) * inv;
        const R t2 = (box.mx[axis] - o[axis]) * inv;
        tmin = smax(tmin, smin(t1, t2));
        tmax = smin(tmax, smax(t1, t2));
    }
    return tmax >= smax(tmin, static_cast<R>(0));
}
// surroundingBox, AABB.cpp:53-60 (argument order matters for signed zeros).
template <class R>
inline AABB<R> surrounding(const AABB<R>& b1, const AABB<R>& b2) {
    return AABB<R>{vmin(b1.mn, b2.mn), vmax(b1.mx, b2.mx)};
}
// AABB::getSurfaceArea, AABB.cpp:37-50.
template <class R>
inline R surface_area(const AABB<R>& b) {
    const R lx = b.mx.x - b.mn.x;
    const R ly = b.mx.y - b.mn.y;
    const R lz = b.mx.z - b.mn.z;
    const R bottomTop = (static_cast<R>(2) * lx) * lz;
    const R sideXY = (static_cast<R>(2) * lx) * ly;
    const R sideZY = (static_cast<R>(2) * lz) * ly;
    return (bottomTop + sideXY) + sideZY;
}

// ----------------------------------------------------------- data types --
template <class R>
struct Material {
    V3<R> Le, Kd, Ks, Kt;
    R ior{1};
};
template <class R>
struct Ray {
    V3<R> o, d;
    int32_t depth{1};
    uint32_t exclude{MRT_ID_NONE};  // Ray::primitive_ as an id
};
// Intersection, app/MobileRT/Intersection.hpp:15-21 (+ u, v for parity output).
template <class R>
struct Isect {
    V3<R> point, normal;
    const Material<R>* material{nullptr};
    R length{static_cast<R>(kRayLengthMax)};
    uint32_t prim{MRT_ID_NONE};   // Intersection::primitive_ (self-exclusion key)
    uint32_t hitId{MRT_ID_NONE};  // category << 30 | index of the hit (reporting)
    int32_t materialId{0};
    V3<R> color{static_cast<R>(1), static_cast<R>(1), static_cast<R>(1)};
    R u{0}, v{0};
};

// Counters kept by the oracle's traversals (interior pairs, primitive tests).
struct TraceCounters {
    uint64_t pairs = 0;
    uint64_t prims = 0;
    std::vector<uint32_t>* cells = nullptr;  // optional: grid cells visited, in walk order
};

// --------------------------------------------------------------- Triangle --
// app/MobileRT/Shapes/Triangle.hpp:14-26, Triangle.cpp.
template <class R>
struct Triangle {
    V3<R> AC, AB, A, nA, nB, nC, cA, cB, cC;
    int32_t materialId{-1};
    uint32_t id{MRT_ID_NONE};

    static Triangle from(const mrt_triangle& t, uint32_t id) {
        Triangle r;
        r.AC = from3<R>(t.ac); r.AB = from3<R>(t.ab); r.A = from3<R>(t.a);
        r.nA = from3<R>(t.na); r.nB = from3<R>(t.nb); r.nC = from3<R>(t.nc);
        r.cA = from3<R>(t.ca); r.cB = from3<R>(t.cb); r.cC = from3<R>(t.cc);
        r.materialId = t.material_id;
        r.id = id;
        return r;
    }

    // Triangle::intersect(Intersection*, Ray), Triangle.cpp:57-103.
    bool intersect(Isect<R>* is, const Ray<R>& ray) const {
        if (ray.exclude == id) return false;
        const V3<R> perp = cross(ray.d, AC);
        const R det = dot(AB, perp);
        if (std::abs(det) < static_cast<R>(kEpsilon)) return false;
        const R inv = static_cast<R>(1) / det;
        const V3<R> s = ray.o - A;
        const R u = inv * dot(s, perp);
        if (u < 0 || u > 1) return false;
        const V3<R> q = cross(s, AB);
        const R v = inv * dot(ray.d, q);
        if (v < 0 || (u + v) > 1) return false;
        const R t = inv * dot(AC, q);
        if (t < static_cast<R>(kEpsilon) || t >= is->length) return false;
        const R w = (static_cast<R>(1) - u) - v;
        const V3<R> n = (nA * w + nB * u) + nC * v;
        const V3<R> c = (cA * w + cB * u) + cC * v;
        const V3<R> p = ray.o + ray.d * t;
        is->point = p;
        is->length = t;
        is->normal = n;
        is->prim = id;
        is->hitId = id;
        is->materialId = materialId;
        is->color = c;
        is->material = nullptr;
        is->u = u;
        is->v = v;
        return true;
    }
    // Triangle::intersect(Ray, dist), Triangle.cpp:105-136.
    bool intersect(const Ray<R>& ray, R dist) const {
        if (ray.exclude == id) return false;
        const V3<R> perp = cross(ray.d, AC);
        const R det = dot(AB, perp);
        if (std::abs(det) < static_cast<R>(kEpsilon)) return false;
        const R inv = static_cast<R>(1) / det;
        const V3<R> s = ray.o - A;
        const R u = inv * dot(s, perp);
        if (u < 0 || u > 1) return false;
        const V3<R> q = cross(s, AB);
        const R v = inv * dot(ray.d, q);
        if (v < 0 || (u + v) > 1) return false;
        const R t = inv * dot(AC, q);
        return t > static_cast<R>(kEpsilon) && t < dist;
    }
    // Triangle::getAABB, Triangle.cpp:145-152.
    AABB<R> aabb() const {
        const V3<R> B = A + AB;
        const V3<R> C = A + AC;
        return AABB<R>{vmin(A, vmin(B, C)), vmax(A, vmax(B, C))};
    }
    // Triangle::intersectBox, Triangle.cpp:154-246.
    bool intersect_box(const AABB<R>& box) const {
        auto rayBox = [&](const V3<R>& orig, const V3<R>& vec) -> bool {
            R t_near = static_cast<R>(std::numeric_limits<float>::min());
            R t_far = static_cast<R>(std::numeric_limits<float>::max());
            for (int a = 0; a < 3; ++a) {
                if (std::abs(vec[a]) < static_cast<R>(kEpsilon)) {
                    if ((orig[a] < box.mn[a]) || ((orig[a] + vec[a]) > box.mx[a])) return false;
                } else {
                    R t1 = (box.mn[a] - orig[a]) / vec[a];
                    R t2 = (box.mx[a] - orig[a]) / vec[a];
                    if (t1 > t2) std::swap(t1, t2);
                    t_near = smax(t1, t_near);
                    t_far = smin(t2, t_far);
                    if ((t_near > t_far) || (t_far < 0)) return false;
                }
            }
            return true;
        };
        auto overTriangle = [&](const V3<R>& vec) -> bool {
            const V3<R> perp = cross(vec, AC);
            const R det = dot(AB, perp);
            return std::abs(det) < static_cast<R>(kEpsilon);
        };
        const V3<R> vec = box.mx - box.mn;
        Ray<R> diag;
        diag.o = box.mn;
        diag.d = vec;
        diag.depth = 1;
        diag.exclude = MRT_ID_NONE;
        const bool iAB = rayBox(A, AB);
        const bool iAC = rayBox(A, AC);
        const V3<R> B = A + AB;
        const V3<R> C = A + AC;
        const bool iBC = rayBox(B, C - B);
        Isect<R> is;
        Triangle self = *this;
        self.id = MRT_ID_NONE - 1;  // the diagonal ray carries no primitive
        const bool iRay = self.intersect(&is, diag);
        const bool inside = overTriangle(vec);
        return iAB || iAC || iBC || iRay || inside;
    }
};

// ----------------------------------------------------------------- Sphere --
// app/MobileRT/Shapes/Sphere.cpp:18-123.
template <class R>
struct Sphere {
    V3<R> center;
    R sqr{};
    int32_t materialId{-1};
    uint32_t id{MRT_ID_NONE};

    static Sphere from(const mrt_sphere& s, uint32_t id) {
        Sphere r;
        r.center = from3<R>(s.center);
        r.sqr = static_cast<R>(s.sq_radius);
        r.materialId = s.material_id;
        r.id = id;
        return r;
    }
    R root(const Ray<R>& ray, bool* ok) const {
        const V3<R> oc = center - ray.o;
        const R proj = dot(oc, ray.d);
        const R ocm = length(oc);
        const R A = dot(ray.d, ray.d);
        const R B = static_cast<R>(2) * -proj;
        const R C = ocm * ocm - sqr;
        const R disc = B * B - (static_cast<R>(4) * A) * C;
        if (disc < 0) { *ok = false; return 0; }
        const R rd = std::sqrt(disc);
        const R t1 = -B + rd;
        const R t2 = -B - rd;
        *ok = true;
        return smin(t1, t2) / (static_cast<R>(2) * A);
    }
    bool intersect(Isect<R>* is, const Ray<R>& ray) const {
        bool ok;
        const R t = root(ray, &ok);
        if (!ok) return false;
        if (t < static_cast<R>(1.0e-05f) || t >= is->length) return false;
        const V3<R> p = ray.o + ray.d * t;
        const V3<R> n = normalize(p - center);
        is->point = p;
        is->length = t;
        is->normal = n;
        is->prim = MRT_ID_NONE;  // Sphere.cpp:50 stores no primitive
        is->hitId = id;
        is->materialId = materialId;
        is->color = V3<R>{1, 1, 1};
        is->material = nullptr;
        is->u = 0;
        is->v = 0;
        return true;
    }
    bool intersect(const Ray<R>& ray, R dist) const {
        bool ok;
        const R t = root(ray, &ok);
        if (!ok) return false;
        return t > static_cast<R>(1.0e-05f) && t < dist;
    }
    AABB<R> aabb() const {
        const R r = std::sqrt(sqr);
        return AABB<R>{center - r, center + r};
    }
    bool intersect_box(const AABB<R>& box) const {
        R dmin = 0;
        for (int a = 0; a < 3; ++a) {
            if (center[a] < box.mn[a]) {
                dmin = dmin + (center[a] - box.mn[a]) * (center[a] - box.mn[a]);
            } else if (center[a] > box.mx[a]) {
                dmin = dmin + (center[a] - box.mx[a]) * (center[a] - box.mx[a]);
            }
        }
        return dmin <= sqr;
    }
};

// ------------------------------------------------------------------ Plane --
// app/MobileRT/Shapes/Plane.cpp:17-155.
template <class R>
struct Plane {
    V3<R> normal, point;
    int32_t materialId{-1};
    uint32_t id{MRT_ID_NONE};

    static Plane from(const mrt_plane& p, uint32_t id) {
        Plane r;
        r.normal = from3<R>(p.normal);
        r.point = from3<R>(p.point);
        r.materialId = p.material_id;
        r.id = id;
        return r;
    }
    bool intersect(Isect<R>* is, const Ray<R>& ray) const {
        if (ray.exclude == id) return false;
        const R np = dot(normal, ray.d);
        if (std::abs(np) < static_cast<R>(kEpsilon)) return false;
        const V3<R> vp = point - ray.o;
        const R sp = dot(normal, vp);
        const R t = sp / np;
        if (t < static_cast<R>(kEpsilon) || t >= is->length) return false;
        is->point = ray.o + ray.d * t;
        is->length = t;
        is->normal = normal;
        is->prim = id;
        is->hitId = id;
        is->materialId = materialId;
        is->color = V3<R>{1, 1, 1};
        is->material = nullptr;
        is->u = 0;
        is->v = 0;
        return true;
    }
    bool intersect(const Ray<R>& ray, R dist) const {
        if (ray.exclude == id) return false;
        const R np = dot(normal, ray.d);
        if (std::abs(np) < static_cast<R>(kEpsilon)) return false;
        const V3<R> vp = point - ray.o;
        const R sp = dot(normal, vp);
        const R t = sp / np;
        return t > static_cast<R>(kEpsilon) && t < dist;
    }
    // Plane::getRightVector / getAABB, Plane.cpp:80-104.
    AABB<R> aabb() const {
        V3<R> right{};
        if (normal.x >= 1) right = V3<R>{0, 1, 1};
        else if (normal.y >= 1) right = V3<R>{1, 0, 1};
        else if (normal.z >= 1) right = V3<R>{1, 1, 0};
        else if (normal.x <= -1) right = V3<R>{0, 1, 1};
        else if (normal.y <= -1) right = V3<R>{1, 0, 1};
        else if (normal.z <= -1) right = V3<R>{1, 1, 0};
        right = normalize(right);
        return AABB<R>{point + right * static_cast<R>(-100.0f), point + right * static_cast<R>(100.0f)};
    }
    // Plane::distance, Plane.cpp:106-121 (left-to-right sums as written).
    R distance(const V3<R>& p) const {
        const R d = (normal.x * -point.x + normal.y * -point.y) + normal.z * -point.z;
        const R num = ((normal.x * p.x + normal.y * p.y) + normal.z * p.z) + d;
        const R den = std::sqrt((normal.x * normal.x + normal.y * normal.y) + normal.z * normal.z);
        return num / den;
    }
    // Plane::intersectBox, Plane.cpp:123-155.
    bool intersect_box(const AABB<R>& box) const {
        const R dP = distance(box.mx);
        const R dN = distance(box.mn);
        return (dP <= 0 && dN >= 0) || (dP >= 0 && dN <= 0);
    }
};

// -------------------------------------------------------------------- BVH --
// app/MobileRT/Accelerators/BVH.hpp.
struct BVHNodeOut {  // BVHNode layout, BVH.hpp:21-25
    float mn[3], mx[3];
    uint32_t indexOffset, numberPrimitives;
};

template <class R, class T>
struct BVH {
    struct Node {
        AABB<R> box;
        uint32_t indexOffset = 0;
        uint32_t numberPrimitives = 0;
    };
    std::vector<Node> boxes;
    std::vector<T> prims;
    uint32_t leafSize = 2;

    // getSplitIndex_SAH, BVH.hpp:274-311 (right_area keeps the reference's
    // off-by-one: right_area[i] covers boxes[i + maxLeafSize ..]).
    static uint32_t split_sah(const std::vector<AABB<R>>& bx, uint32_t L) {
        const uint32_t N = static_cast<uint32_t>(bx.size());
        std::vector<R> left(N - L, 0), right(N - L, 0);
        AABB<R> lb = bx[0];
        left[0] = surface_area(lb);
        for (uint32_t i = 1; i < N - L; ++i) {
            lb = surrounding(lb, bx[i]);
            left[i] = surface_area(lb);
        }
        AABB<R> rb = bx[N - 1];
        right[N - L - 1] = surface_area(rb);
        for (uint32_t i = N - 2; i > L; --i) {
            rb = surrounding(rb, bx[i]);
            right[i - L] = surface_area(rb);
        }
        uint32_t split = L;
        R minSah = static_cast<R>(L) * left[L - 1] + static_cast<R>(N - L) * right[L - 1];
        for (uint32_t i = L; i < N - L; ++i) {
            const R sl = static_cast<R>(i + 1) * left[i];
            const R sr = static_cast<R>(N - (i + 1)) * right[i];
            const R sah = sl + sr;
            if (sah < minSah) {
                split = i + 1;
                minSah = sah;
            }
        }
        return split;
    }

    // BVH<T>::BVH + build, BVH.hpp:64-155 (explicit stacks; the reference's
    // 512-entry arrays become vectors that may grow).
    void build(std::vector<T>&& p, uint32_t leaf) {
        leafSize = leaf;
        boxes.clear();
        if (p.empty()) {
            boxes.push_back(Node{});
            prims.clear();
            return;
        }
        prims = std::move(p);
        const uint32_t n = static_cast<uint32_t>(prims.size());
        boxes.resize(static_cast<size_t>(n) * 2 - 1);
        uint32_t id = 0, begin = 0, end = n, maxId = 0;
        std::vector<uint32_t> sId{0}, sBegin{0}, sEnd{0};
        do {
            Node& node = boxes[id];
            node.box = prims[begin].aabb();
            std::vector<AABB<R>> bx{node.box};
            const uint32_t count = end - begin;
            bx.reserve(count);
            for (uint32_t i = begin + 1; i < end; ++i) {
                const AABB<R> nb = prims[i].aabb();
                node.box = surrounding(nb, node.box);
                bx.push_back(nb);
            }
            if (count <= leafSize) {
                node.indexOffset = begin;
                node.numberPrimitives = count;
                id = sId.back(); sId.pop_back();
                begin = sBegin.back(); sBegin.pop_back();
                end = sEnd.back(); sEnd.pop_back();
            } else {
                const uint32_t left = maxId + 1;
                const uint32_t right = left + 1;
                node.indexOffset = left;
                maxId = std::max(right, maxId);
                const uint32_t split = count <= 2 * leafSize ? leafSize : split_sah(bx, leafSize);
                sId.push_back(right);
                sBegin.push_back(begin + split);
                sEnd.push_back(end);
                id = left;
                end = begin + split;
            }
        } while (!sId.empty());
        boxes.resize(static_cast<size_t>(maxId) + 1);
    }

    // BVH<T>::trace, BVH.hpp:158-213.
    bool trace(Isect<R>* is, const Ray<R>& ray, TraceCounters* tc) const {
        if (prims.empty()) return false;
        uint32_t id = 0;
        std::vector<uint32_t> stack{0};
        bool hit = false;
        do {
            const Node& node = boxes[id];
            if (slab(node.box, ray.o, ray.d)) {
                if (node.numberPrimitives > 0) {
                    for (uint32_t i = 0; i < node.numberPrimitives; ++i) {
                        if (tc) ++tc->prims;
                        hit |= prims[node.indexOffset + i].intersect(is, ray);
                    }
                    id = stack.back(); stack.pop_back();
                } else {
                    if (tc) ++tc->pairs;
                    const uint32_t l = node.indexOffset, r = l + 1;
                    const bool tl = slab(boxes[l].box, ray.o, ray.d);
                    const bool tr = slab(boxes[r].box, ray.o, ray.d);
                    if (!tl && !tr) {
                        id = stack.back(); stack.pop_back();
                    } else {
                        id = tl ? l : r;
                        if (tl && tr) stack.push_back(r);
                    }
                }
            } else {
                id = stack.back(); stack.pop_back();
            }
        } while (!stack.empty());
        return hit;
    }
    // BVH<T>::shadowTrace, BVH.hpp:216-271.
    bool shadow(const Ray<R>& ray, R dist, TraceCounters* tc) const {
        if (prims.empty()) return false;
        uint32_t id = 0;
        std::vector<uint32_t> stack{0};
        do {
            const Node& node = boxes[id];
            if (slab(node.box, ray.o, ray.d)) {
                if (node.numberPrimitives > 0) {
                    for (uint32_t i = 0; i < node.numberPrimitives; ++i) {
                        if (tc) ++tc->prims;
                        if (prims[node.indexOffset + i].intersect(ray, dist)) return true;
                    }
                    id = stack.back(); stack.pop_back();
                } else {
                    if (tc) ++tc->pairs;
                    const uint32_t l = node.indexOffset, r = l + 1;
                    const bool tl = slab(boxes[l].box, ray.o, ray.d);
                    const bool tr = slab(boxes[r].box, ray.o, ray.d);
                    if (!tl && !tr) {
                        id = stack.back(); stack.pop_back();
                    } else {
                        id = tl ? l : r;
                        if (tl && tr) stack.push_back(r);
                    }
                }
            } else {
                id = stack.back(); stack.pop_back();
            }
        } while (!stack.empty());
        return false;
    }
};

// Near-first walk of the same tree with culling by the nearest hit so far
// (closest hit) or an exit at the first blocker (any hit).  Measurement only:
// it counts the interior visits (each fetches the node's pair of children,
// two 32-byte BVHNodes) and primitive tests an ordered traversal of the
// reference tree needs, the implementation-independent traffic of a ray; it
// returns nothing the renderer uses.
struct OrderedCounts {
    std::atomic<uint64_t> rays_closest{0}, pairs_closest{0}, tests_closest{0};
    std::atomic<uint64_t> rays_shadow{0}, pairs_shadow{0}, tests_shadow{0};
    std::atomic<uint64_t> pairs_lights{0}, tests_lights{0};
};

template <class R, class T>
bool ordered_walk(const BVH<R, T>& b, const Ray<R>& ray, Isect<R>* is, bool any, R dist, uint64_t* pairs,
                  uint64_t* tests) {
    if (b.prims.empty()) return false;
    const V3<R> inv{static_cast<R>(1) / ray.d.x, static_cast<R>(1) / ray.d.y, static_cast<R>(1) / ray.d.z};
    auto entry = [&](const AABB<R>& bx, R limit, R* e) {
        R t0 = 0, t1 = limit;
        for (int a = 0; a < 3; ++a) {
            R n = (bx.mn[a] - ray.o[a]) * inv[a], f = (bx.mx[a] - ray.o[a]) * inv[a];
            if (n > f) std::swap(n, f);
            if (!(n <= t1) || !(f >= t0)) return false;  // NaN (0 * inf) keeps the box
            t0 = std::max(t0, n);
            t1 = std::min(t1, f);
        }
        *e = t0;
        return t0 <= t1;
    };
    auto limit = [&]() { return any ? dist : is->length; };
    R e0;
    if (!entry(b.boxes[0].box, limit(), &e0)) return false;
    std::vector<std::pair<uint32_t, R>> stack{{0u, e0}};
    bool hit = false;
    while (!stack.empty()) {
        const auto [id, e] = stack.back();
        stack.pop_back();
        if (e > limit()) continue;
        const auto& node = b.boxes[id];
        if (node.numberPrimitives > 0) {
            for (uint32_t i = 0; i < node.numberPrimitives; ++i) {
                ++*tests;
                if (any) {
                    if (b.prims[node.indexOffset + i].intersect(ray, dist)) return true;
                } else {
                    hit |= b.prims[node.indexOffset + i].intersect(is, ray);
                }
            }
            continue;
        }
        ++*pairs;
        const uint32_t l = node.indexOffset, r = l + 1;
        R el, er;
        const bool hl = entry(b.boxes[l].box, limit(), &el), hr = entry(b.boxes[r].box, limit(), &er);
        if (hl && hr) {
            const bool lFirst = el <= er;
            stack.emplace_back(lFirst ? r : l, lFirst ? er : el);
            stack.emplace_back(lFirst ? l : r, lFirst ? el : er);
        } else if (hl) {
            stack.emplace_back(l, el);
        } else if (hr) {
            stack.emplace_back(r, er);
        }
    }
    return hit;
}

// ------------------------------------------------------------------ Naive --
// app/MobileRT/Accelerators/Naive.hpp:55-74.
template <class R, class T>
struct Naive {
    std::vector<T> prims;
    bool trace(Isect<R>* is, const Ray<R>& ray, TraceCounters* tc) const {
        bool hit = false;
        for (const T& p : prims) {
            if (tc) ++tc->prims;
            hit |= p.intersect(is, ray);
        }
        return hit;
    }
    bool shadow(const Ray<R>& ray, R dist, TraceCounters* tc) const {
        for (const T& p : prims) {
            if (tc) ++tc->prims;
            if (p.intersect(ray, dist)) return true;
        }
        return false;
    }
};

// ----------------------------------------------------------- RegularGrid --
// app/MobileRT/Accelerators/RegularGrid.hpp.
inline uint32_t bit_counter(uint32_t n) {  // Utils.cpp:26-33
    uint32_t c = 0;
    while (n > 0) { ++c; n >>= 1; }
    return c;
}
inline uint32_t round_up_pow2(uint32_t v) {  // Utils.cpp:15-24
    --v;
    v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
    ++v;
    return v;
}
// Grid size, Shader.cpp:76-108.
inline int32_t grid_size_for(uint32_t n) {
    const int32_t aux = n > 6 ? static_cast<int32_t>(round_up_pow2(bit_counter(n))) : 2;
    const int32_t aux2 = static_cast<int32_t>(round_up_pow2(bit_counter(static_cast<uint32_t>(aux))) / 2);
    return std::max(aux * aux2, aux);
}

template <class R, class T>
struct RegularGrid {
    std::vector<std::vector<uint32_t>> cells;  // indices into prims
    std::vector<T> prims;
    int32_t gridSize = 0;
    uint32_t gridShift = 0;
    AABB<R> ext;
    V3<R> SR, CW;
    bool hasPrims = false;

    // RegularGrid ctor + addPrimitives, RegularGrid.hpp:50-154.
    void build(const AABB<R>& bounds, std::vector<T>&& p, int32_t gs) {
        gridSize = gs;
        gridShift = bit_counter(static_cast<uint32_t>(gs)) - 1;
        ext = bounds;
        const V3<R> e = ext.mx - ext.mn;
        SR = V3<R>{static_cast<R>(gs) / e.x, static_cast<R>(gs) / e.y, static_cast<R>(gs) / e.z};
        CW = e * (static_cast<R>(1) / static_cast<R>(gs));
        hasPrims = !p.empty();
        prims = std::move(p);
        cells.assign(static_cast<size_t>(gs) * gs * gs, {});
        const R sizeX = ext.mx.x - ext.mn.x, sizeY = ext.mx.y - ext.mn.y, sizeZ = ext.mx.z - ext.mn.z;
        const R dx = sizeX / static_cast<R>(gs), dy = sizeY / static_cast<R>(gs), dz = sizeZ / static_cast<R>(gs);
        const R dxr = dx > 0 ? static_cast<R>(1) / dx : static_cast<R>(1);
        const R dyr = dy > 0 ? static_cast<R>(1) / dy : static_cast<R>(1);
        const R dzr = dz > 0 ? static_cast<R>(1) / dz : static_cast<R>(1);
        for (uint32_t pi = 0; pi < prims.size(); ++pi) {
            const AABB<R> b = prims[pi].aabb();
            int32_t x1 = static_cast<int32_t>((b.mn.x - ext.mn.x) * dxr);
            int32_t x2 = static_cast<int32_t>((b.mx.x - ext.mn.x) * dxr) + 1;
            x1 = std::max(0, x1);
            x2 = std::min(gs - 1, x2);
            x2 = std::abs(sizeX) < static_cast<R>(kEpsilon) ? 0 : x2;
            x1 = std::min(x1, x2);
            int32_t y1 = static_cast<int32_t>((b.mn.y - ext.mn.y) * dyr);
            int32_t y2 = static_cast<int32_t>((b.mx.y - ext.mn.y) * dyr) + 1;
            y1 = std::max(0, y1);
            y2 = std::min(gs - 1, y2);
            y2 = std::abs(sizeY) < static_cast<R>(kEpsilon) ? 0 : y2;
            y1 = std::min(y2, y1);
            int32_t z1 = static_cast<int32_t>((b.mn.z - ext.mn.z) * dzr);
            int32_t z2 = static_cast<int32_t>((b.mx.z - ext.mn.z) * dzr) + 1;
            z1 = std::max(0, z1);
            z2 = std::min(gs - 1, z2);
            z2 = std::abs(sizeZ) < static_cast<R>(kEpsilon) ? 0 : z2;
            z1 = std::min(z2, z1);
            for (int32_t x = x1; x <= x2; ++x)
                for (int32_t y = y1; y <= y2; ++y)
                    for (int32_t z = z1; z <= z2; ++z) {
                        const size_t idx = static_cast<size_t>(x) + static_cast<size_t>(y) * gs +
                                           static_cast<size_t>(z) * gs * gs;
                        const V3<R> pos{ext.mn.x + static_cast<R>(x) * dx, ext.mn.y + static_cast<R>(y) * dy,
                                        ext.mn.z + static_cast<R>(z) * dz};
                        const AABB<R> cell{pos, pos + V3<R>{dx, dy, dz}};
                        if (prims[pi].intersect_box(cell)) cells[idx].push_back(pi);
                    }
        }
    }

    struct Dda {
        int32_t X, Y, Z, stepX, stepY, stepZ, outX, outY, outZ;
        V3<R> tmax, tdelta;
    };
    void setup(const Ray<R>& ray, Dda* g) const {
        const V3<R> c = (ray.o - ext.mn) * SR;
        g->X = std::min(gridSize - 1, std::max(0, static_cast<int32_t>(c.x)));
        g->Y = std::min(gridSize - 1, std::max(0, static_cast<int32_t>(c.y)));
        g->Z = std::min(gridSize - 1, std::max(0, static_cast<int32_t>(c.z)));
        V3<R> cb{};
        int32_t* P[3] = {&g->X, &g->Y, &g->Z};
        int32_t* S[3] = {&g->stepX, &g->stepY, &g->stepZ};
        int32_t* O[3] = {&g->outX, &g->outY, &g->outZ};
        for (int a = 0; a < 3; ++a) {
            if (ray.d[a] > 0) {
                *S[a] = 1;
                *O[a] = gridSize;
                cb[a] = ext.mn[a] + static_cast<R>(*P[a] + 1) * CW[a];
            } else {
                *S[a] = -1;
                *O[a] = -1;
                cb[a] = ext.mn[a] + static_cast<R>(*P[a]) * CW[a];
            }
        }
        g->tmax = V3<R>{};
        g->tdelta = V3<R>{};
        for (int a = 0; a < 3; ++a) {
            if (std::abs(ray.d[a]) > static_cast<R>(kEpsilon)) {
                const R rr = static_cast<R>(1) / ray.d[a];
                g->tmax[a] = (cb[a] - ray.o[a]) * rr;
                g->tdelta[a] = (CW[a] * static_cast<R>(*S[a])) * rr;
            } else {
                g->tmax[a] = static_cast<R>(kRayLengthMax);
            }
        }
    }
    uint32_t cell_index(const Dda& g) const {
        return static_cast<uint32_t>(g.X) + (static_cast<uint32_t>(g.Y) << gridShift) +
               (static_cast<uint32_t>(g.Z) << (gridShift * 2u));
    }

    // RegularGrid<T>::trace, RegularGrid.hpp:157-337 (phase 1 until any hit,
    // then phase 2 until the nearest t is closer than the next cell wall).
    bool trace(Isect<R>* is, const Ray<R>& ray, TraceCounters* tc) const {
        if (!hasPrims) return false;
        Dda g;
        setup(ray, &g);
        bool hit = false;
        bool phase2 = false;
        while (!phase2) {
            if (tc && tc->cells) tc->cells->push_back(cell_index(g));
            const std::vector<uint32_t> list = cells[cell_index(g)];  // copied as in :243
            for (uint32_t pi : list) {
                if (tc) ++tc->prims;
                hit |= prims[pi].intersect(is, ray);
                if (hit) { phase2 = true; break; }
            }
            if (phase2) break;
            if (tc) ++tc->pairs;
            if (g.tmax.x < g.tmax.y) {
                if (g.tmax.x < g.tmax.z) {
                    g.X += g.stepX;
                    if (g.X == g.outX) return hit;
                    g.tmax.x = g.tmax.x + g.tdelta.x;
                } else {
                    g.Z += g.stepZ;
                    if (g.Z == g.outZ) return hit;
                    g.tmax.z = g.tmax.z + g.tdelta.z;
                }
            } else {
                if (g.tmax.y < g.tmax.z) {
                    g.Y += g.stepY;
                    if (g.Y == g.outY) return hit;
                    g.tmax.y = g.tmax.y + g.tdelta.y;
                } else {
                    g.Z += g.stepZ;
                    if (g.Z == g.outZ) return hit;
                    g.tmax.z = g.tmax.z + g.tdelta.z;
                }
            }
        }
        while (true) {
            // Phase 2 starts in the cell phase 1 ended in (logged there).
            if (tc && tc->cells && tc->cells->back() != cell_index(g)) tc->cells->push_back(cell_index(g));
            const std::vector<uint32_t> list = cells[cell_index(g)];  // copied as in :290
            for (uint32_t pi : list) {
                if (tc) ++tc->prims;
                hit |= prims[pi].intersect(is, ray);
            }
            if (tc) ++tc->pairs;
            if (g.tmax.x < g.tmax.y) {
                if (g.tmax.x < g.tmax.z) {
                    if (is->length < g.tmax.x) break;
                    g.X += g.stepX;
                    if (g.X == g.outX) break;
                    g.tmax.x = g.tmax.x + g.tdelta.x;
                } else {
                    if (is->length < g.tmax.z) break;
                    g.Z += g.stepZ;
                    if (g.Z == g.outZ) break;
                    g.tmax.z = g.tmax.z + g.tdelta.z;
                }
            } else {
                if (g.tmax.y < g.tmax.z) {
                    if (is->length < g.tmax.y) break;
                    g.Y += g.stepY;
                    if (g.Y == g.outY) break;
                    g.tmax.y = g.tmax.y + g.tdelta.y;
                } else {
                    if (is->length < g.tmax.z) break;
                    g.Z += g.stepZ;
                    if (g.Z == g.outZ) break;
                    g.tmax.z = g.tmax.z + g.tdelta.z;
                }
            }
        }
        return hit;
    }
    // RegularGrid<T>::shadowTrace, RegularGrid.hpp:340-465.
    bool shadow(const Ray<R>& ray, R dist, TraceCounters* tc) const {
        if (!hasPrims) return false;
        Dda g;
        setup(ray, &g);
        while (true) {
            const std::vector<uint32_t> list = cells[cell_index(g)];  // copied as in :425
            for (uint32_t pi : list) {
                if (tc) ++tc->prims;
                if (prims[pi].intersect(ray, dist)) return true;
            }
            if (tc) ++tc->pairs;
            if (g.tmax.x < g.tmax.y) {
                if (g.tmax.x < g.tmax.z) {
                    g.X += g.stepX;
                    if (g.X == g.outX) return false;
                    g.tmax.x = g.tmax.x + g.tdelta.x;
                } else {
                    g.Z += g.stepZ;
                    if (g.Z == g.outZ) return false;
                    g.tmax.z = g.tmax.z + g.tdelta.z;
                }
            } else {
                if (g.tmax.y < g.tmax.z) {
                    g.Y += g.stepY;
                    if (g.Y == g.outY) return false;
                    g.tmax.y = g.tmax.y + g.tdelta.y;
                } else {
                    g.Z += g.stepZ;
                    if (g.Z == g.outZ) return false;
                    g.tmax.z = g.tmax.z + g.tdelta.z;
                }
            }
        }
    }
};

}  // namespace orc
