// Host-side acceleration-structure builders of the product library.
#pragma once
#include <cstdint>
#include <vector>

#include "mrt_scene.hpp"

namespace mrt {

// BVHNode of app/MobileRT/Accelerators/BVH.hpp:21-25 (box + offset + count).
struct RefNode {
    Box box;
    uint32_t offset;
    uint32_t count;
};

// Output of a BVH build in the device layout.
struct BvhOut {
    std::vector<RefNode> ref;      // reference node array (for checks)
    std::vector<BNode> pairs;      // 64-byte child pairs, root pair first
    Box root;
    uint32_t root_ref = 0, root_count = 0;
    uint32_t depth = 0;
};

// The reference builder: BVH<T>::BVH + build + getSplitIndex_SAH
// (BVH.hpp:64-155, 274-311) over primitive boxes given in scene order.
void build_bvh_reference(const std::vector<Box>& boxes, uint32_t leaf_size, BvhOut* out);
// The child pairs, root and depth of out->ref (a reference node array, as
// the builder or the GPU build leaves it).
void bvh_pairs_from_ref(BvhOut* out);

// Primitive boxes (getAABB of Triangle.cpp:145-152, Sphere.cpp:91-96,
// Plane.cpp:80-104).
Box triangle_box(const mrt_triangle& t);
Box sphere_box(const mrt_sphere& s);
Box plane_box(const mrt_plane& p);

// Triangle::intersectBox (Triangle.cpp:154-246), Sphere::intersectBox
// (Sphere.cpp:102-123), Plane::intersectBox (Plane.cpp:123-155).
bool triangle_intersect_box(const mrt_triangle& t, const Box& box);
bool sphere_intersect_box(const mrt_sphere& s, const Box& box);
bool plane_intersect_box(const mrt_plane& p, const Box& box);

// RegularGrid in CSR form.
struct GridOut {
    int32_t grid_size = 0;
    uint32_t grid_shift = 0;
    Box ext;
    f3 sr, cw;
    std::vector<uint32_t> offsets;  // grid_size^3 + 1
    std::vector<uint32_t> refs;
};
// Grid size of Shader.cpp:76-108.
int32_t grid_size_for(uint32_t n);
// RegularGrid ctor + addPrimitives, RegularGrid.hpp:50-154.  `bounds` is the
// scene box of Shader.cpp:52-62 (primitives + camera, widened by Epsilon).
template <class BoxFn, class HitFn>
void build_grid(uint32_t n, const Box& bounds, int32_t grid_size, BoxFn box_of, HitFn intersects, GridOut* out);

// Conservative BVH over light triangles (any builder will do: the light
// query keeps the (t, index) key, so only reachability matters; leaf boxes
// are widened so that the slab test never rejects a point of the triangle).
void build_light_bvh(const std::vector<Box>& boxes, BvhOut* out, std::vector<uint32_t>* perm);

// One child slot of a 4-wide node: its box, a reference (a wide node for an
// interior slot, the first primitive slot for a leaf) and the leaf word
// (0 for an interior slot).
struct WideSlot {
    Box box;
    uint32_t ref, word;
};

// Writes wide node `index` (8 float4, SoA boxes + (ref, word) pairs, unused
// slots all-NaN) into `out`.
void put_wide_node(std::vector<float4>* out, uint32_t index, const WideSlot* slots, int n);

// 4-wide BVH over leaf units (box + leaf word each, every unit kept as its
// own leaf slot) by binned surface-area heuristic on the unit centres.  Nodes
// get indices from `base` on and are written into `out` from its start (node
// `base` first); the tree's top is returned as at most `max_top` slots for the
// caller's root nodes, and *depth receives the number of wide levels below
// them.
void build_wide_sah(const std::vector<WideSlot>& units, uint32_t base, int max_top, std::vector<float4>* out,
                    std::vector<WideSlot>* top, uint32_t* depth);

}  // namespace mrt

#include "mrt_build_grid.inl"
