// Device-side scene layout (HBM) shared by the kernels and the host uploader.
//
// Triangles are split into a hot and a cold array:
//   tri_geo   3 x float4 per triangle  A.xyz AB.x | AB.yz AC.xy | AC.z id 0 0
//             the 36 bytes Triangle::intersect reads (Triangle.cpp:57-103)
//             plus the primitive id, padded to 48 so a triangle is three
//             aligned 16-byte loads; after the scene triangles come the area
//             light triangles in the order the light tree's leaves use
//             (slots of the combined walk tree below);
//   tri_shade 5 x float4 per triangle  nA nB nC cA cB cC materialId, read
//             once per closest hit for shading.
// BVH nodes are 64-byte child pairs: the two child boxes of one reference
// interior node (BVH.hpp:186-192 tests both children together) plus the two
// child references, so one interior visit is one 64-byte fetch:
//   q0 = L.min.xyz L.max.x, q1 = L.max.yz R.min.xy, q2 = R.min.z R.max.xyz,
//   q3 = (Lref, Lcount, Rref, Rcount); count > 0 marks a leaf whose
//   primitives are [ref, ref + count), count == 0 an interior pair index.
// The combined walk tree (persistent kernels) uses the same pairs with one
// change: a leaf child's count word holds (count << 27) | first slot, the
// slot indexing tri_geo, so queuing a leaf is one word (leaves hold at most
// 16 primitives; 27 slot bits let one arena hold the trees of many scenes).  Pair 0 holds the
// triangle root box and the light root box, pair 1 the triangle root box and
// an all-NaN box that no ray enters; triangle pairs start at pair 2, light
// pairs at walk_light_pair.
#pragma once
#include "mrt_math.hpp"
#include "../../include/mobilert_gpu.h"

namespace mrt {

struct alignas(16) BNode {
    float4 q0, q1, q2;
    uint4 q3;
};
static_assert(sizeof(BNode) == 64, "BVH node pair must be 64 bytes");

// Walk tree: root pairs and the leaf word (count << kSlotBits | slot).
// The wide walk tree replaces every interior child of a pair by that child's
// own two children: 4 child slots per 128-byte node, structure of arrays
//   w[0..2] = min.x, min.y, min.z of slots 0..3, w[3..5] = max.x/y/z,
//   w[6] = (ref0, word0, ref1, word1), w[7] = (ref2, word2, ref3, word3),
// word != 0 a leaf word, word == 0 with a finite box an interior slot whose
// ref is a wide node, unused slots all-NaN.  Wide nodes 0 and 1 are the two
// roots; nodes from light pairs follow those from triangle pairs.
constexpr uint32_t kWalkClosestRoot = 0u, kWalkAnyRoot = 1u, kWalkFirstPair = 2u;
constexpr uint32_t kSlotBits = 27u;
constexpr uint32_t kSlotMask = (1u << kSlotBits) - 1u;

// One accelerator (Shader::{naive,regularGrid,bvh}{Planes,Spheres,Triangles}_).
struct Accel {
    uint32_t count;         // primitives in this category
    int32_t kind;           // MRT_NAIVE / MRT_REGULAR_GRID / MRT_BVH
    // BVH
    const BNode* nodes;     // interior pairs; the root pair is nodes[0]
    const uint32_t* perm;   // leaf slot -> primitive index (NULL: identity)
    float root_mn[3], root_mx[3];
    uint32_t root_ref, root_count;  // root_count > 0: the root is a leaf
    uint32_t depth;         // longest root-to-leaf interior path
    // RegularGrid, app/MobileRT/Accelerators/RegularGrid.hpp:18-24
    const uint32_t* cell_off;  // CSR offsets, grid_size^3 + 1
    const uint32_t* cell_ref;  // primitive indices, ascending per cell
    int32_t grid_size;
    uint32_t grid_shift;
    float ext_mn[3], sr[3], cw[3];
};

struct LightInfo {
    int32_t kind;           // MRT_LIGHT_AREA / MRT_LIGHT_POINT
    uint32_t material;      // index of the radiance material
    float pos[3];           // PointLight position
    mrt_sampler sampler;    // AreaLight point sampler
};

struct SceneDev {
    const float4* tri_geo;
    const float4* tri_shade;
    const float4* sph;       // center.xyz, sq_radius
    const int32_t* sph_mat;
    const float4* pla;       // 2 per plane: normal.xyz matId-bits, point.xyz 0
    const float4* mat;       // 4 per material: Le.xyz Kd.x | Kd.yz Ks.xy | Ks.z Kt.xyz | ior
    const float4* light_geo;   // 3 per light (same layout as tri_geo)
    const float4* light_shade; // 5 per light
    const LightInfo* light_info;
    uint32_t n_lights;
    uint32_t n_area_lights;
    uint32_t lights_linear;  // 1: Scene::traceLights as a linear scan
    uint32_t has_branches;   // any material with Ks or Kt > 0
    Accel acc[3];            // planes, spheres, triangles (Shader.cpp:144-148 order)
    Accel light_acc;         // conservative BVH over area lights (lights_linear == 0)
    const BNode* walk_nodes; // combined walk tree (NULL: not built)
    uint32_t walk_light_pair;  // first light pair of the walk tree
    uint32_t walk_light_slot;  // first light slot of tri_geo (= triangle count)
    const float4* wide_nodes;  // wide walk tree, 8 float4 per node
    uint32_t wide_light_node;  // first wide node from a light pair
    const uint4* cw_nodes;     // compressed wide walk tree (mrt_cwt.hpp), 4 uint4 per node
    const float4* cw_recs;     // its leaf records, 8 float4 each
};

// Material fields.
struct MatV {
    f3 le, kd, ks, kt;
    float ior;
};

}  // namespace mrt
