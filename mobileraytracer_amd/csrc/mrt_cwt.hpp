// Compressed wide walk tree: the wide walk tree (mrt_scene.hpp) with its
// child boxes quantized to 8 bits per bound, and each leaf's triangles stored
// beside the leaf's exact box in one 128-byte record.  For rays whose inverse
// direction is finite in every axis the reference slab test is monotone under
// box inclusion, so interior boxes may be any boxes that contain the exact
// ones: a quantized box is rounded outward, a ray that passes a leaf's exact
// box passes every quantized box above it, and the leaf test applies the
// reference predicate to the exact box (BVH.hpp:186-192 on the reference's
// own leaf box) before it tests the triangles.  The leaves a ray reaches, and
// so the (t, id) answer, are those of the exact wide tree.
//
// Node (64 bytes, 4 x uint4), slots k = 0..3:
//   n0 = origin.x, origin.y, origin.z (float bits), ex | ey << 8 | ez << 16 | kinds << 24
//   n1 = qmin.x, qmin.y, qmin.z, qmax.x   (slot k in byte k of each word)
//   n2 = qmax.y, qmax.z, ref0, ref1
//   n3 = ref2, ref3, 0, 0
// kind (bits 2k of kinds): empty, interior (ref = node), leaf (ref = leaf
// word), open (a box with an infinite or NaN bound, entered at distance 0).
// Slot bounds on axis a: origin_a + q * 2^(e_a - 127), rounded like every
// float sum here (no contraction), and the builder lowers qmin / raises qmax
// until those floats contain the exact bounds.
//
// Leaf record r (128 bytes, 8 float4), for the leaf word (count << 27 | s):
// records s .. s + ceil(count / 2) - 1 (a leaf's first slot is its first
// record, and leaves never share slots), record s + j holding the leaf's exact
// box and its triangles s + 2j and s + 2j + 1 in the tri_geo layout:
//   r0 = min.xyz, max.x   r1 = max.yz, 0, 0
//   r2..r4 = triangle 2j, r5..r7 = triangle 2j + 1 (when the leaf has it)
#pragma once
#include <cstdint>

#include "mrt_math.hpp"

namespace mrt {

constexpr uint32_t kCwEmpty = 0u, kCwInterior = 1u, kCwLeaf = 2u, kCwOpen = 3u;

MRT_HD float cw_float(uint32_t u) {
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}
MRT_HD uint32_t cw_bits(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return u;
}

// 2^(e - 127) for a biased exponent e in [1, 254].
MRT_HD float cw_scale(uint32_t e) { return cw_float(e << 23); }

// A quantized bound as the walk computes it.
MRT_HD float cw_deq(float origin, uint32_t q, float scale) { return origin + static_cast<float>(q) * scale; }

MRT_HD uint32_t cw_byte(uint32_t word, uint32_t k) { return (word >> (8u * k)) & 0xFFu; }

}  // namespace mrt
