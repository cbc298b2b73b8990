// Vector math of the ray-scene path, shared by the gfx950 kernels and the host
// builders.  Every operation is written with an explicit evaluation order and
// the library is compiled with -ffp-contract=off, so a*b+c is two rounded
// operations on the GPU and on the host alike.  The order follows glm's
// generic formulas, which the reference calls through (glm::dot, cross,
// normalize, length, reflect, refract, min, max; see
// app/MobileRT/Shapes/Triangle.cpp:57-103 and app/MobileRT/Shader.cpp:193-223):
//   dot(a,b)     = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b)   = (a.y*b.z - b.y*a.z, a.z*b.x - b.z*a.x, a.x*b.y - b.x*a.y)
//   normalize(v) = v * (1 / sqrt(dot(v,v)))      (IEEE divide and sqrt)
//   reflect(I,N) = I - (N * dot(N,I)) * 2
//   refract      = k = 1 - (e*e)*(1 - d*d); k >= 0 ? e*I - (e*d + sqrt(k))*N : 0
//   smin(a,b)    = (b < a) ? b : a,  smax(a,b) = (a < b) ? b : a  (std::min/max)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MRT_HD __host__ __device__ __forceinline__

namespace mrt {

constexpr float kRayLengthMax = 1.0e+30f;  // RayLengthMax, app/MobileRT/Utils.hpp:33
constexpr float kEpsilon = 1.0e-06f;       // Epsilon, app/MobileRT/Utils.hpp:37
constexpr uint32_t kBlocks = 256;          // NumberOfBlocks, app/MobileRT/Utils.hpp:36

struct f3 {
    float x, y, z;
};

MRT_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
MRT_HD f3 ld3(const float* p) { return f3{p[0], p[1], p[2]}; }
MRT_HD void st3(float* p, f3 v) { p[0] = v.x; p[1] = v.y; p[2] = v.z; }
MRT_HD f3 add(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
MRT_HD f3 sub(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
MRT_HD f3 mul(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
MRT_HD f3 muls(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
MRT_HD f3 divs(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
MRT_HD f3 adds(f3 a, float s) { return f3{a.x + s, a.y + s, a.z + s}; }
MRT_HD f3 subs(f3 a, float s) { return f3{a.x - s, a.y - s, a.z - s}; }
MRT_HD float comp(f3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

MRT_HD float dot(f3 a, f3 b) {
    const float xx = a.x * b.x;
    const float yy = a.y * b.y;
    const float zz = a.z * b.z;
    return (xx + yy) + zz;
}
MRT_HD f3 cross(f3 a, f3 b) {
    const float x = a.y * b.z - b.y * a.z;
    const float y = a.z * b.x - b.z * a.x;
    const float z = a.x * b.y - b.x * a.y;
    return f3{x, y, z};
}
MRT_HD float length(f3 v) { return sqrtf(dot(v, v)); }
MRT_HD f3 normalize(f3 v) {
    const float inv = 1.0f / sqrtf(dot(v, v));
    return muls(v, inv);
}
MRT_HD f3 reflect(f3 i, f3 n) { return sub(i, muls(muls(n, dot(n, i)), 2.0f)); }
MRT_HD f3 refract(f3 i, f3 n, float eta) {
    const float d = dot(n, i);
    const float k = 1.0f - (eta * eta) * (1.0f - d * d);
    if (k >= 0.0f) {
        const float f = eta * d + sqrtf(k);
        return sub(muls(i, eta), muls(n, f));
    }
    return f3{0.0f, 0.0f, 0.0f};
}
MRT_HD float smin(float a, float b) { return (b < a) ? b : a; }
MRT_HD float smax(float a, float b) { return (a < b) ? b : a; }
MRT_HD f3 vmin(f3 a, f3 b) { return f3{smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)}; }
MRT_HD f3 vmax(f3 a, f3 b) { return f3{smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)}; }
MRT_HD bool any_gt0(f3 v) { return v.x > 0.0f || v.y > 0.0f || v.z > 0.0f; }

struct Box {
    f3 mn, mx;
};
// surroundingBox, app/MobileRT/Boxes/AABB.cpp:53-60 (argument order kept:
// it decides which zero survives when +0 and -0 meet).
MRT_HD Box surrounding(const Box& b1, const Box& b2) { return Box{vmin(b1.mn, b2.mn), vmax(b1.mx, b2.mx)}; }
// AABB::getSurfaceArea, AABB.cpp:37-50.
MRT_HD float surface_area(const Box& b) {
    const float lx = b.mx.x - b.mn.x;
    const float ly = b.mx.y - b.mn.y;
    const float lz = b.mx.z - b.mn.z;
    const float bt = (2.0f * lx) * lz;
    const float xy = (2.0f * lx) * ly;
    const float zy = (2.0f * lz) * ly;
    return (bt + xy) + zy;
}

// Ray-box slab test of ::MobileRT::intersect(AABB, Ray), AABB.cpp:15-35, with
// the inverse direction precomputed (the reference divides per call; the
// quotient 1/d is the same value).  Returns hit and the entry distance
// max(tmin, 0) for ordering and culling.
MRT_HD bool slab(const f3& mn, const f3& mx, const f3& o, const f3& inv, float* entry) {
    const float t1x = (mn.x - o.x) * inv.x;
    const float t2x = (mx.x - o.x) * inv.x;
    float tmin = smin(t1x, t2x);
    float tmax = smax(t1x, t2x);
    const float t1y = (mn.y - o.y) * inv.y;
    const float t2y = (mx.y - o.y) * inv.y;
    tmin = smax(tmin, smin(t1y, t2y));
    tmax = smin(tmax, smax(t1y, t2y));
    const float t1z = (mn.z - o.z) * inv.z;
    const float t2z = (mx.z - o.z) * inv.z;
    tmin = smax(tmin, smin(t1z, t2z));
    tmax = smin(tmax, smax(t1z, t2z));
    const float e = smax(tmin, 0.0f);
    *entry = e;
    return tmax >= e;
}

}  // namespace mrt
