// Device-side trees of a GPU-built triangle BVH: from the reference node array
// (left on the device by gpu_build_bvh) the child pairs of the closest-hit
// kernels, the triangle part of the combined walk tree and its 4-wide walk
// nodes, written in place with no host copy of the tree.  The layouts are
// those of mrt_scene.hpp; the wide nodes are the host collapse's (each
// interior pair at even depth below the root, its interior children
// replaced by their two children), numbered in pair order instead of the
// host's depth-first order (numbering changes no traversal result).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mrt_build.hpp"

namespace mrt {

// Device scratch of the tree passes, grown on demand and kept between builds.
struct WalkGpuScratch {
    void* p = nullptr;
    size_t bytes = 0;
    WalkGpuScratch() = default;
    WalkGpuScratch(const WalkGpuScratch&) = delete;
    WalkGpuScratch& operator=(const WalkGpuScratch&) = delete;
    ~WalkGpuScratch();
};

// What the host needs of the triangle trees: sizes, depths and the top.
struct TriTreeInfo {
    uint32_t pairs;        // interior reference nodes (= child pairs)
    uint32_t wide;         // wide nodes (interior pairs at even depth)
    uint32_t depth;        // longest chain of interior pairs, the root pair 1
    uint32_t wide_levels;  // level of the deepest wide node (top slots at level 2), 1 if none
    RefNode root;          // reference node 0 (interior)
    RefNode child[2];      // its children
    uint32_t child_wide[2];  // wide index (from 0) of each interior child
};

// Pass 1 over the cnt reference nodes at `ref` (device memory, ref[0]
// interior, every interior node's children at offset and offset + 1 < cnt):
// interior ranks, parent links, pair depths and the wide numbering, kept in
// the scratch for pass 2; fills *info.  Synchronizes the stream.
hipError_t gpu_tri_tree_plan(const RefNode* ref, uint32_t cnt, WalkGpuScratch* sc, TriTreeInfo* info,
                             hipStream_t st);

// Pass 2: pairs[info.pairs] (child pairs: interior children by rank, leaves
// as (first, count)); walk[info.pairs] (the same pairs with interior
// references + pair_base and leaves as leaf words (count << 27 | first));
// wide[8 * info.wide] (wide node i at wide + 8 i, interior slots referring to
// wide_base + their index).  Any of the three may be null.
hipError_t gpu_tri_tree_write(const RefNode* ref, uint32_t cnt, const WalkGpuScratch& sc, const TriTreeInfo& info,
                              uint32_t pair_base, uint32_t wide_base, BNode* pairs, BNode* walk, float4* wide,
                              hipStream_t st);

// Copies of walk trees with their references rebased (a render server's
// scene arena): compressed wide nodes (4 uint4 each) with leaf words +
// tri_base and interior references + wide_base; walk pairs with leaf
// references and words + tri_base and interior references (count 0, finite
// box) + walk_base.  src and dst may not overlap.
hipError_t gpu_rebase_cw(const uint4* src, size_t n, uint4* dst, uint32_t tri_base, uint32_t wide_base,
                         hipStream_t st);
hipError_t gpu_rebase_walk(const BNode* src, size_t n, BNode* dst, uint32_t tri_base, uint32_t walk_base,
                           hipStream_t st);

// The device records of n scene triangles from their mrt_triangle array
// (device memory): tri_geo (3 float4: A.xyz AB.x | AB.yz AC.xy | AC.z id 0 0,
// id = MRT_MAKE_ID(MRT_CAT_TRIANGLE, index)) and tri_shade (5 float4: nA nB
// nC cA cB cC materialId), the same bits as the host packing.
hipError_t gpu_pack_triangles(const mrt_triangle* tris, uint32_t n, float4* geo, float4* shade, hipStream_t st);

// The compressed wide walk tree (mrt_cwt.hpp) of the first n_wide nodes of a
// wide walk tree: n_wide nodes of 4 uint4, and the leaf records (8 float4
// each, indexed by a leaf's first slot) of every leaf the nodes reference;
// tri_geo holds n_slots slots.
hipError_t gpu_compress_walk(const float4* wide, uint32_t n_wide, const float4* tri_geo, uint32_t n_slots,
                             uint4* nodes, float4* recs, hipStream_t st);

// Boxes (6 floats: min xyz, max xyz) of the first n triangles of a tri_geo
// array (3 float4 per triangle: A.xyz AB.x | AB.yz AC.xy | AC.z id), the
// same arithmetic as the host's triangle_box.
hipError_t gpu_tri_boxes(const float4* tri_geo, uint32_t n, float* boxes6, hipStream_t st);

}  // namespace mrt
