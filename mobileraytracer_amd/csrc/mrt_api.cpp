// C ABI of the gfx950 ray-scene library: contexts, scene upload, accelerator
// builds, traversal / shading entry points and the tile renderer.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "mrt_internal.hpp"

namespace mrt {
thread_local std::string g_error;
}  // namespace mrt

namespace {

using namespace mrt;

uint64_t buffers_bytes(const mrt_scene* s) {
    uint64_t b = s->tri_geo.bytes + s->tri_shade.bytes + s->sph.bytes + s->sph_mat.bytes + s->pla.bytes +
                 s->mat.bytes + s->light_geo.bytes + s->light_shade.bytes + s->light_info.bytes;
    for (int i = 0; i < 4; ++i) b += s->nodes[i].bytes + s->perm[i].bytes;
    for (int i = 0; i < 3; ++i) b += s->cell_off[i].bytes + s->cell_ref[i].bytes;
    return b + s->walk_nodes.bytes + s->wide_nodes.bytes + s->cw_nodes.bytes + s->cw_recs.bytes;
}

// The compressed wide walk tree (mrt_cwt.hpp) of the scene's n_wide wide
// nodes, written on the device from the wide tree and tri_geo.
int build_compressed_walk(mrt_scene* s, uint32_t n_wide, hipStream_t st) {
    s->dev.cw_nodes = nullptr;
    s->dev.cw_recs = nullptr;
    const uint64_t slots = s->tris.size() + static_cast<uint64_t>(s->n_light_slots);
    if (n_wide == 0u || slots == 0u) return MRT_OK;
    const uint64_t bytes = 64ull * n_wide + 128ull * slots;
    mrt_ctx* ctx = s->ctx;
    if (ctx->budget && buffers_bytes(s) + bytes > ctx->budget) return MRT_OK;  // the exact wide tree serves
    int rc;
    if ((rc = s->cw_nodes.alloc(64ull * n_wide)) != MRT_OK) return rc;
    if ((rc = s->cw_recs.alloc(128ull * slots)) != MRT_OK) return rc;
    // Records no leaf starts at stay zero (never read; zero triangles hit nothing).
    MRT_HIP(hipMemsetAsync(s->cw_recs.p, 0, 128ull * slots, st));
    MRT_HIP(gpu_compress_walk(s->wide_nodes.as<float4>(), n_wide, s->tri_geo.as<float4>(),
                              static_cast<uint32_t>(slots), s->cw_nodes.as<uint4>(), s->cw_recs.as<float4>(), st));
    s->dev.cw_nodes = s->cw_nodes.as<uint4>();
    s->dev.cw_recs = s->cw_recs.as<float4>();
    return MRT_OK;
}

void pack_geo(const mrt_triangle& t, uint32_t id, float4* g) {
    float idf;
    std::memcpy(&idf, &id, 4);
    g[0] = make_float4(t.a[0], t.a[1], t.a[2], t.ab[0]);
    g[1] = make_float4(t.ab[1], t.ab[2], t.ac[0], t.ac[1]);
    g[2] = make_float4(t.ac[2], idf, 0.0f, 0.0f);
}
void pack_shade(const mrt_triangle& t, float4* s) {
    float mid;
    std::memcpy(&mid, &t.material_id, 4);
    s[0] = make_float4(t.na[0], t.na[1], t.na[2], t.nb[0]);
    s[1] = make_float4(t.nb[1], t.nb[2], t.nc[0], t.nc[1]);
    s[2] = make_float4(t.nc[2], t.ca[0], t.ca[1], t.ca[2]);
    s[3] = make_float4(t.cb[0], t.cb[1], t.cb[2], t.cc[0]);
    s[4] = make_float4(t.cc[1], t.cc[2], mid, 0.0f);
}

void fill_accel_bvh(Accel* a, const BvhOut& b, const DevBuf& nodes, const DevBuf* perm, uint32_t count) {
    a->count = count;
    a->kind = MRT_BVH;
    a->nodes = nodes.as<BNode>();
    a->perm = perm ? perm->as<uint32_t>() : nullptr;
    st3(a->root_mn, b.root.mn);
    st3(a->root_mx, b.root.mx);
    a->root_ref = b.root_ref;
    a->root_count = b.root_count;
    a->depth = b.depth;
}

int upload_bvh(mrt_scene* s, int slot, const BvhOut& b, hipStream_t st) {
    if (b.pairs.empty()) return MRT_OK;
    return s->nodes[slot].upload(b.pairs.data(), b.pairs.size() * sizeof(BNode), st);
}

// Child `side` (0 left, 1 right) of a walk-tree pair.
void set_child(BNode* p, int side, const Box& b, uint32_t ref, uint32_t word) {
    if (side == 0) {
        p->q0 = make_float4(b.mn.x, b.mn.y, b.mn.z, b.mx.x);
        p->q1.x = b.mx.y;
        p->q1.y = b.mx.z;
        p->q3.x = ref;
        p->q3.y = word;
    } else {
        p->q1.z = b.mn.x;
        p->q1.w = b.mn.y;
        p->q2 = make_float4(b.mn.z, b.mx.x, b.mx.y, b.mx.z);
        p->q3.z = ref;
        p->q3.w = word;
    }
}

uint32_t leaf_word(uint32_t slot, uint32_t count) { return (count << kSlotBits) | slot; }

// Copies the pairs of one tree into the walk tree: interior references move
// by `pair_base`, leaves become (count << 27 | slot_base + first).
void append_pairs(std::vector<BNode>* w, const BvhOut& t, uint32_t pair_base, uint32_t slot_base) {
    for (const BNode& p : t.pairs) {
        BNode q = p;
        q.q3.x = p.q3.y ? slot_base + p.q3.x : pair_base + p.q3.x;
        q.q3.y = p.q3.y ? leaf_word(slot_base + p.q3.x, p.q3.y) : 0u;
        q.q3.z = p.q3.w ? slot_base + p.q3.z : pair_base + p.q3.z;
        q.q3.w = p.q3.w ? leaf_word(slot_base + p.q3.z, p.q3.w) : 0u;
        w->push_back(q);
    }
}

// The root child of a tree: its root box, as a leaf or as its first pair.
void set_root_child(BNode* p, int side, const BvhOut& t, uint32_t pair_base, uint32_t slot_base) {
    if (t.root_count > 0u) set_child(p, side, t.root, slot_base + t.root_ref, leaf_word(slot_base + t.root_ref, t.root_count));
    else set_child(p, side, t.root, pair_base, 0u);
}

bool is_nan_box(const BNode& p, int side) {
    return side == 0 ? std::isnan(p.q0.x) : std::isnan(p.q1.z);
}

WideSlot pair_child(const BNode& p, int side) {
    if (side == 0) return {Box{f3{p.q0.x, p.q0.y, p.q0.z}, f3{p.q0.w, p.q1.x, p.q1.y}}, p.q3.x, p.q3.y};
    return {Box{f3{p.q1.z, p.q1.w, p.q2.x}, f3{p.q2.y, p.q2.z, p.q2.w}}, p.q3.z, p.q3.w};
}

// Walk-tree pairs by index: [0, lo_n) from lo and, when hi is set, [hi_base,
// hi_base + hi_n) from hi (the root pairs and the light pairs of a walk tree
// whose triangle pairs live on the device only).
struct PairView {
    const BNode* lo;
    uint32_t lo_n;
    const BNode* hi = nullptr;
    uint32_t hi_base = 0, hi_n = 0;
    const BNode& operator[](uint32_t i) const { return i < lo_n ? lo[i] : hi[i - hi_base]; }
    uint32_t size() const { return hi ? hi_base + hi_n : lo_n; }
};

// Child slots of the wide node made from pair `pi`: each leaf or empty child
// as it is, each interior child replaced by its two children.
int wide_slots(const PairView& w, uint32_t pi, WideSlot* out) {
    int n = 0;
    for (int side = 0; side < 2; ++side) {
        const WideSlot c = pair_child(w[pi], side);
        if (c.word != 0u || is_nan_box(w[pi], side)) {
            out[n++] = c;
        } else {
            out[n++] = pair_child(w[c.ref], 0);
            out[n++] = pair_child(w[c.ref], 1);
        }
    }
    return n;
}

// Depth-first collapse of the pairs below `top` into wide nodes numbered
// from `first` (each a pair with its interior children expanded), written to
// `nodes` from its start; top slots sit at level 2, below the roots, and
// *max_level receives the deepest level reached.  `index` maps pairs to
// their wide nodes (kNone: not yet numbered).
uint32_t collapse_pairs(const PairView& w, std::vector<WideSlot>* top, uint32_t first, std::vector<float4>* nodes,
                        std::vector<uint32_t>* index, uint32_t* max_level) {
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    uint32_t next = first;
    std::vector<std::pair<uint32_t, uint32_t>> todo;  // (pair, level)
    auto map_slot = [&](WideSlot c, uint32_t level) {
        if (c.word != 0u || std::isnan(c.box.mn.x)) return c;
        if ((*index)[c.ref] == kNone) {
            (*index)[c.ref] = next++;
            todo.emplace_back(c.ref, level);
        }
        c.ref = (*index)[c.ref];
        return c;
    };
    for (WideSlot& c : *top) c = map_slot(c, 2u);
    while (!todo.empty()) {
        const auto [pi, level] = todo.back();
        todo.pop_back();
        *max_level = std::max(*max_level, level);
        WideSlot sl[4];
        const int n = wide_slots(w, pi, sl);
        for (int k = 0; k < n; ++k) sl[k] = map_slot(sl[k], level + 1u);
        put_wide_node(nodes, (*index)[pi] - first, sl, n);
    }
    return next - first;
}

// The light slots of root 0: the light root's children (interior root), the
// root itself (a leaf: the linear light scan), or none.
std::vector<WideSlot> light_top(const PairView& w) {
    std::vector<WideSlot> top;
    if (!is_nan_box(w[kWalkClosestRoot], 1)) {
        const WideSlot c = pair_child(w[kWalkClosestRoot], 1);
        if (c.word != 0u) {
            top.push_back(c);
        } else {
            top.push_back(pair_child(w[c.ref], 0));
            top.push_back(pair_child(w[c.ref], 1));
        }
    }
    return top;
}

// Builds the wide walk tree (layout in mrt_scene.hpp) for rays whose inverse
// direction is finite in every axis.  For such rays the reference slab test
// is monotone under box inclusion (the products (bound - o) * inv round
// monotonically), so a leaf whose own box the ray passes also passes every box
// that contains it: the leaves a ray reaches in the reference tree are exactly
// those whose own boxes it passes, whatever tree is built above them, as long
// as every interior box is the exact union of the leaf boxes below it.  The
// triangle part is therefore a binned-SAH 4-wide tree over the reference
// BVH's leaves (each kept with its own box), or (collapse) the reference
// pairs collapsed; the light part collapses the light tree's pairs (each
// interior child replaced by its two children).
// Nodes: 0 and 1 the roots, then the triangle nodes, then the light nodes.
int build_wide_tree(mrt_scene* s, const std::vector<BNode>& wv, const BvhOut& tri, hipStream_t st, bool collapse) {
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    const PairView w{wv.data(), static_cast<uint32_t>(wv.size())};
    std::vector<WideSlot> lightTop = light_top(w);
    std::vector<uint32_t> index(w.size(), kNone);
    uint32_t maxDepth = 1u;
    std::vector<float4> triNodes, out(16, make_float4(0, 0, 0, 0));
    std::vector<WideSlot> triTop;
    uint32_t nTriNodes = 0;
    if (collapse) {
        // Triangle top: the triangle root of root pair 0, expanded when
        // interior; then its pairs collapsed.
        if (!is_nan_box(w[kWalkClosestRoot], 0)) {
            const WideSlot c = pair_child(w[kWalkClosestRoot], 0);
            if (c.word != 0u) {
                triTop.push_back(c);
            } else {
                triTop.push_back(pair_child(w[c.ref], 0));
                triTop.push_back(pair_child(w[c.ref], 1));
            }
        }
        nTriNodes = collapse_pairs(w, &triTop, kWalkFirstPair, &triNodes, &index, &maxDepth);
    } else {
        // Triangle units: the reference leaves, under a binned-SAH tree.
        std::vector<WideSlot> units;
        if (!s->tris.empty()) {
            if (tri.root_count > 0) {
                units.push_back(WideSlot{tri.root, tri.root_ref, leaf_word(tri.root_ref, tri.root_count)});
            } else {
                for (const BNode& p : tri.pairs) {
                    for (int side = 0; side < 2; ++side) {
                        const WideSlot c = pair_child(p, side);
                        if (c.word != 0u) units.push_back(WideSlot{c.box, c.ref, leaf_word(c.ref, c.word)});
                    }
                }
            }
        }
        uint32_t triDepth = 0;
        build_wide_sah(units, kWalkFirstPair, 4 - static_cast<int>(lightTop.size()), &triNodes, &triTop, &triDepth);
        nTriNodes = static_cast<uint32_t>(triNodes.size() / 8);
        maxDepth = 1u + triDepth;
    }
    const uint32_t lightNode = kWalkFirstPair + nTriNodes;
    std::vector<float4> lightNodes;
    collapse_pairs(w, &lightTop, lightNode, &lightNodes, &index, &maxDepth);
    // Roots: 0 = triangle top + light top, 1 = triangle top (shadow rays).
    WideSlot r0[4], r1[4];
    int n0 = 0, n1 = 0;
    for (const WideSlot& c : triTop) r0[n0++] = r1[n1++] = c;
    for (const WideSlot& c : lightTop) r0[n0++] = c;
    put_wide_node(&out, kWalkClosestRoot, r0, n0);
    put_wide_node(&out, kWalkAnyRoot, r1, n1);
    out.resize(8ull * kWalkFirstPair);
    out.insert(out.end(), triNodes.begin(), triNodes.end());
    out.insert(out.end(), lightNodes.begin(), lightNodes.end());
    int rc;
    if ((rc = s->wide_nodes.upload(out.data(), out.size() * sizeof(float4), st)) != MRT_OK) return rc;
    MRT_HIP(hipStreamSynchronize(st));  // `out` is released on return
    s->dev.wide_nodes = s->wide_nodes.as<float4>();
    s->dev.wide_light_node = lightNode;
    if ((rc = build_compressed_walk(s, static_cast<uint32_t>(out.size() / 8), st)) != MRT_OK) return rc;
    // A wide visit pushes at most three children.
    s->depth_need = std::max(s->depth_need, 3u * maxDepth + 1u);
    return MRT_OK;
}

// The root pairs of a walk tree (triangle root `tri` at pair_base, lights
// after it) and whether the tree's slots and leaves fit the leaf word.
bool walk_roots(mrt_scene* s, const BvhOut& tri, uint32_t lightBase, BNode w[kWalkFirstPair]) {
    const uint64_t slots = s->tris.size() + static_cast<uint64_t>(s->n_light_slots);
    if (slots >= (1ull << kSlotBits)) return false;
    const float nan = std::numeric_limits<float>::quiet_NaN();
    const float inf = std::numeric_limits<float>::infinity();
    const Box empty{f3{nan, nan, nan}, f3{nan, nan, nan}};
    const uint32_t nTri = static_cast<uint32_t>(s->tris.size());
    std::memset(w, 0, kWalkFirstPair * sizeof(BNode));
    for (uint32_t r = 0; r < kWalkFirstPair; ++r) {
        if (nTri > 0u) set_root_child(&w[r], 0, tri, kWalkFirstPair, 0u);
        else set_child(&w[r], 0, empty, 0u, 0u);
        set_child(&w[r], 1, empty, 0u, 0u);
    }
    if (s->n_light_slots > 0u) {
        if (s->dev.lights_linear) set_child(&w[kWalkClosestRoot], 1, Box{f3{-inf, -inf, -inf}, f3{inf, inf, inf}}, nTri,
                                            leaf_word(nTri, s->n_light_slots));
        else set_root_child(&w[kWalkClosestRoot], 1, s->light_bvh, lightBase, nTri);
    }
    return true;
}

uint32_t light_max_leaf(const mrt_scene* s) {
    uint32_t m = s->n_light_slots <= kLinearLights ? s->n_light_slots : 0u;
    for (const BNode& p : s->light_bvh.pairs) m = std::max({m, p.q3.y, p.q3.w});
    return m;
}

// Builds the combined walk tree of the persistent kernels (layout in
// mrt_scene.hpp).  Scenes whose slots or leaves do not fit the 27/5-bit leaf
// word get none; they run the non-persistent kernels.
int build_walk_tree(mrt_scene* s, const BvhOut& tri, hipStream_t st, bool collapse) {
    s->dev.walk_nodes = nullptr;
    uint32_t maxLeaf = std::max(tri.root_count, light_max_leaf(s));
    for (const BNode& p : tri.pairs) maxLeaf = std::max({maxLeaf, p.q3.y, p.q3.w});
    if (maxLeaf >= (1u << (32u - kSlotBits))) return MRT_OK;
    const uint32_t nTri = static_cast<uint32_t>(s->tris.size());
    const uint32_t triBase = kWalkFirstPair;
    const uint32_t lightBase = triBase + static_cast<uint32_t>(tri.pairs.size());
    std::vector<BNode> w(kWalkFirstPair);
    if (!walk_roots(s, tri, lightBase, w.data())) return MRT_OK;
    append_pairs(&w, tri, triBase, 0u);
    append_pairs(&w, s->light_bvh, lightBase, nTri);
    int rc;
    if ((rc = s->walk_nodes.upload(w.data(), w.size() * sizeof(BNode), st)) != MRT_OK) return rc;
    s->dev.walk_nodes = s->walk_nodes.as<BNode>();
    s->dev.walk_light_pair = lightBase;
    s->dev.walk_light_slot = nTri;
    s->depth_need = std::max(s->depth_need, std::max(tri.depth, s->light_bvh.depth) + 1u);
    if ((rc = build_wide_tree(s, w, tri, st, collapse)) != MRT_OK) return rc;
    MRT_HIP(hipStreamSynchronize(st));
    return MRT_OK;
}

// The triangle BVH of a scene built on the GPU and kept there: boxes from the
// uploaded triangles, the reference node array (gpu_build_bvh), then the
// child pairs, the walk tree's triangle pairs and its wide nodes written by
// the device (mrt_walk_gpu.hip); the host adds the root pairs, the light
// pairs and the light part of the wide tree (the collapse of build_wide_tree
// in both parts, the triangle wide nodes numbered in pair order).  *resident
// stays false when the GPU build declines (*b untouched) or the root is a
// leaf (*b then holds the host form of the one-node tree); the caller goes
// on with the host steps.  Callers hold ctx->mu.
int build_tri_bvh_resident(mrt_scene* s, uint32_t leaf, hipStream_t st, BvhOut* b, bool* built, bool* resident) {
    *built = false;
    *resident = false;
    mrt_ctx* ctx = s->ctx;
    const uint32_t n = static_cast<uint32_t>(s->tris.size());
    if (n == 0) return MRT_OK;
    int rc;
    if ((rc = ctx->bvh_boxes.alloc(static_cast<size_t>(n) * sizeof(Box))) != MRT_OK) return rc;
    MRT_HIP(gpu_tri_boxes(s->tri_geo.as<float4>(), n, ctx->bvh_boxes.as<float>(), st));
    if ((rc = ctx->bvh_nodes.alloc((2 * static_cast<size_t>(n) - 1) * sizeof(RefNode))) != MRT_OK) return rc;
    uint32_t cnt = 0;
    const RefNode* ref = ctx->bvh_nodes.as<RefNode>();
    const hipError_t e =
        gpu_build_bvh(ctx->bvh_boxes.as<float>(), n, leaf, ctx->bvh_nodes.as<RefNode>(), &cnt, &ctx->bvh_scratch, st);
    if (e == hipErrorNotSupported) return MRT_OK;
    MRT_HIP(e);
    *built = true;
    TriTreeInfo info{};
    // hipErrorInvalidValue: a leaf root or a tree the device passes decline;
    // any other error is reported.
    const hipError_t pe = cnt < 3 ? hipErrorInvalidValue : gpu_tri_tree_plan(ref, cnt, &ctx->walk_scratch, &info, st);
    if (pe != hipSuccess && pe != hipErrorInvalidValue) MRT_HIP(pe);
    if (pe != hipSuccess) {
        // The host steps.
        b->ref.resize(cnt);
        MRT_HIP(hipMemcpyAsync(b->ref.data(), ref, cnt * sizeof(RefNode), hipMemcpyDeviceToHost, st));
        MRT_HIP(hipStreamSynchronize(st));
        bvh_pairs_from_ref(b);
        return MRT_OK;
    }
    b->ref.clear();
    b->pairs.clear();
    b->root = info.root.box;
    b->root_ref = 0;
    b->root_count = 0;
    b->depth = info.depth;
    const uint64_t need = buffers_bytes(s) + 2ull * info.pairs * sizeof(BNode);
    if (ctx->budget && need > ctx->budget) return fail(MRT_E_BUDGET, "BVH exceeds the HBM budget");
    const int c = MRT_CAT_TRIANGLE;
    if ((rc = s->nodes[c].alloc(static_cast<size_t>(info.pairs) * sizeof(BNode))) != MRT_OK) return rc;
    fill_accel_bvh(&s->dev.acc[c], *b, s->nodes[c], nullptr, n);
    s->depth_need = std::max(s->depth_need, b->depth);
    *resident = true;

    // Walk tree: roots, triangle pairs (device), light pairs.
    s->dev.walk_nodes = nullptr;
    const uint32_t nTri = n;
    const uint32_t lightBase = kWalkFirstPair + info.pairs;
    BNode roots[kWalkFirstPair];
    const bool walk = leaf < (1u << (32u - kSlotBits)) && light_max_leaf(s) < (1u << (32u - kSlotBits)) &&
                      walk_roots(s, *b, lightBase, roots);
    if (!walk) {
        MRT_HIP(gpu_tri_tree_write(ref, cnt, ctx->walk_scratch, info, kWalkFirstPair, kWalkFirstPair,
                                   s->nodes[c].as<BNode>(), nullptr, nullptr, st));
        MRT_HIP(hipStreamSynchronize(st));
        return MRT_OK;
    }
    std::vector<BNode> lp;
    append_pairs(&lp, s->light_bvh, lightBase, nTri);
    const size_t nWalk = kWalkFirstPair + static_cast<size_t>(info.pairs) + lp.size();
    if ((rc = s->walk_nodes.alloc(nWalk * sizeof(BNode))) != MRT_OK) return rc;
    BNode* wdev = s->walk_nodes.as<BNode>();
    MRT_HIP(hipMemcpyAsync(wdev, roots, sizeof(roots), hipMemcpyHostToDevice, st));
    if (!lp.empty())
        MRT_HIP(hipMemcpyAsync(wdev + lightBase, lp.data(), lp.size() * sizeof(BNode), hipMemcpyHostToDevice, st));

    // Wide tree: roots, triangle nodes (device), light nodes (host collapse).
    const PairView w{roots, kWalkFirstPair, lp.data(), lightBase, static_cast<uint32_t>(lp.size())};
    std::vector<WideSlot> lightTop = light_top(w);
    std::vector<WideSlot> triTop;
    for (int k = 0; k < 2; ++k) {
        const RefNode& ch = info.child[k];
        if (ch.count != 0u) triTop.push_back(WideSlot{ch.box, ch.offset, leaf_word(ch.offset, ch.count)});
        else triTop.push_back(WideSlot{ch.box, kWalkFirstPair + info.child_wide[k], 0u});
    }
    const uint32_t lightNode = kWalkFirstPair + info.wide;
    std::vector<uint32_t> index(w.size(), 0xFFFFFFFFu);
    uint32_t maxDepth = info.wide_levels;
    std::vector<float4> lightNodes, top(16, make_float4(0, 0, 0, 0));
    collapse_pairs(w, &lightTop, lightNode, &lightNodes, &index, &maxDepth);
    WideSlot r0[4], r1[4];
    int n0 = 0, n1 = 0;
    for (const WideSlot& t : triTop) r0[n0++] = r1[n1++] = t;
    for (const WideSlot& t : lightTop) r0[n0++] = t;
    put_wide_node(&top, kWalkClosestRoot, r0, n0);
    put_wide_node(&top, kWalkAnyRoot, r1, n1);
    const size_t nWide = 8ull * (kWalkFirstPair + info.wide) + lightNodes.size();
    if ((rc = s->wide_nodes.alloc(nWide * sizeof(float4))) != MRT_OK) return rc;
    float4* wide = s->wide_nodes.as<float4>();
    MRT_HIP(hipMemcpyAsync(wide, top.data(), 8ull * kWalkFirstPair * sizeof(float4), hipMemcpyHostToDevice, st));
    if (!lightNodes.empty())
        MRT_HIP(hipMemcpyAsync(wide + 8ull * lightNode, lightNodes.data(), lightNodes.size() * sizeof(float4),
                               hipMemcpyHostToDevice, st));
    MRT_HIP(gpu_tri_tree_write(ref, cnt, ctx->walk_scratch, info, kWalkFirstPair, kWalkFirstPair,
                               s->nodes[c].as<BNode>(), wdev + kWalkFirstPair, wide + 8ull * kWalkFirstPair, st));
    MRT_HIP(hipStreamSynchronize(st));  // the host arrays are released on return
    s->dev.walk_nodes = wdev;
    s->dev.walk_light_pair = lightBase;
    s->dev.walk_light_slot = nTri;
    s->depth_need = std::max(s->depth_need, std::max(b->depth, s->light_bvh.depth) + 1u);
    s->dev.wide_nodes = wide;
    s->dev.wide_light_node = lightNode;
    s->depth_need = std::max(s->depth_need, 3u * maxDepth + 1u);
    if ((rc = build_compressed_walk(s, static_cast<uint32_t>(nWide / 8), st)) != MRT_OK) return rc;
    MRT_HIP(hipStreamSynchronize(st));
    return MRT_OK;
}

// Render parameters common to mrt_render and mrt_shade_rays.
// Tiles of the 16x16 grid not skipped by tile_skip.
uint32_t rendered_tiles(const RenderParams& p) {
    uint32_t n = 0;
    for (int k = 0; k < 8; ++k) n += static_cast<uint32_t>(__builtin_popcount(~p.tile_skip[k]));
    return n;
}

RenderParams params_from(const mrt_render_desc& d) {
    RenderParams p{};
    p.shader = d.shader;
    p.width = d.width;
    p.height = d.height;
    p.spp = d.samples_pixel;
    p.samples_light = d.samples_light;
    p.depth_min = d.ray_depth_min;
    p.depth_max = d.ray_depth_max;
    p.cam = d.camera;
    p.s_pixel = d.sampler_pixel;
    p.s_rr = d.sampler_rr;
    p.s_light = d.sampler_light;
    p.s_hemi = d.sampler_hemi;
    for (int k = 0; k < 3; ++k) p.max_point[k] = d.depth_max_point[k];
    for (int k = 0; k < 8; ++k) p.tile_skip[k] = d.tile_skip[k];
    return p;
}

int check_sampler(const mrt_sampler& s, bool pixel, const char* what) {
    const bool compat = s.kind == MRT_SAMPLER_HALTON_COMPAT || s.kind == MRT_SAMPLER_STRATIFIED_COMPAT;
    if (s.kind < MRT_SAMPLER_CONSTANT || s.kind > MRT_SAMPLER_STRATIFIED_COMPAT || (compat && !pixel))
        return fail(MRT_E_INVALID, std::string("invalid sampler kind for ") + what);
    if (compat && s.domain_size == 0) return fail(MRT_E_INVALID, std::string("compat sampler needs domain_size: ") + what);
    return MRT_OK;
}

int check_desc(const mrt_scene* s, const mrt_render_desc* d) {
    if (!d) return fail(MRT_E_INVALID, "null render descriptor");
    if (d->shader < MRT_SHADER_NOSHADOWS || d->shader > MRT_SHADER_DIFFUSE)
        return fail(MRT_E_INVALID, "unknown shader " + std::to_string(d->shader));
    if (d->ray_depth_max < 0 || d->ray_depth_max > 15) return fail(MRT_E_UNSUPPORTED, "ray_depth_max must be 0..15");
    if (d->samples_light > 255) return fail(MRT_E_UNSUPPORTED, "samples_light must be <= 255");
    int rc;
    if ((rc = check_sampler(d->sampler_pixel, true, "pixel")) != MRT_OK) return rc;
    if ((rc = check_sampler(d->sampler_rr, false, "russian roulette")) != MRT_OK) return rc;
    if ((rc = check_sampler(d->sampler_light, false, "light choice")) != MRT_OK) return rc;
    if ((rc = check_sampler(d->sampler_hemi, false, "hemisphere")) != MRT_OK) return rc;
    for (const mrt_light& l : s->lights)
        if ((rc = check_sampler(l.sampler, false, "area light")) != MRT_OK) return rc;
    if (d->camera.kind != MRT_CAMERA_PERSPECTIVE && d->camera.kind != MRT_CAMERA_ORTHOGRAPHIC)
        return fail(MRT_E_INVALID, "camera kind");
    return MRT_OK;
}

// Grows the spill area to cover `lanes` lanes; fills the stack config.
int ensure_spill(mrt_scene* s, uint32_t lanes, StackCfg* sc) {
    const uint32_t need = s->depth_need + 1u;
    sc->spill = nullptr;
    sc->spill_stride = lanes;
    sc->cap_spill = 0;
    if (need <= kStackLds) return MRT_OK;
    sc->cap_spill = need - kStackLds;
    const size_t bytes = static_cast<size_t>(sc->cap_spill) * lanes * sizeof(uint32_t);
    const int rc = s->spill.alloc(bytes);
    if (rc != MRT_OK) return rc;
    sc->spill = s->spill.as<uint32_t>();
    return MRT_OK;
}

int prepare_aux(mrt_scene* s) {
    int rc;
    if ((rc = s->counters.alloc(16 * sizeof(unsigned long long))) != MRT_OK) return rc;
    if ((rc = s->err.alloc(sizeof(uint32_t))) != MRT_OK) return rc;
    MRT_HIP(hipMemsetAsync(s->counters.p, 0, 16 * sizeof(unsigned long long), s->ctx->stream));
    MRT_HIP(hipMemsetAsync(s->err.p, 0, sizeof(uint32_t), s->ctx->stream));
    return MRT_OK;
}

int check_err(mrt_scene* s) {
    uint32_t e = 0;
    MRT_HIP(hipMemcpyAsync(&e, s->err.p, sizeof(e), hipMemcpyDeviceToHost, s->ctx->stream));
    MRT_HIP(hipStreamSynchronize(s->ctx->stream));
    if (e & 1u) return fail(MRT_E_INVALID, "traversal stack overflow (BVH deeper than the stack budget)");
    if (e & 2u) return fail(MRT_E_UNSUPPORTED, "too many pending reflection/transmission rays");
    return MRT_OK;
}

// Copies `bytes` from host or device memory into a staging buffer (host) or
// returns the device pointer as is.
int stage(DevBuf& buf, const void* src, size_t bytes, int mem, hipStream_t st, const void** out) {
    if (mem == MRT_MEM_DEVICE) {
        *out = src;
        return MRT_OK;
    }
    const int rc = buf.upload(src, bytes, st);
    if (rc != MRT_OK) return rc;
    *out = buf.p;
    return MRT_OK;
}

uint32_t cap_blocks(const mrt_ctx* c) { return (c->num_cu ? c->num_cu : 256u) * 16u; }

// The wavefront frame.  Each slot group runs its own iteration loop on its
// own stream; the host issues the groups' iterations interleaved, in batches
// of kBatch.  After each batch a group's counters are copied to pinned memory
// and the host reads the previous batch's copy, so the GPU never waits for the
// host.  A group is done when the extension, shadow and ended queues a batch
// left behind are all empty (the batch queued meanwhile finds nothing to do).
// A stop request lets running paths finish without starting new samples.
int run_wavefront(mrt_scene* s, const mrt_render_desc* d, const RenderParams& p, std::vector<WfBuffers>& wbs,
                  bool persistent, uint32_t trace_blocks, uint32_t persist_blocks, const std::vector<StackCfg>& scs,
                  uint32_t* dbm, float* dacc, float* daov, bool* stopped, float* kernel_ms, float* trace_ms,
                  float* busy_ms, uint32_t* launches) {
    constexpr uint32_t kBatch = 4;
    const uint32_t G = static_cast<uint32_t>(wbs.size());
    hipStream_t st[kWfMaxGroups];
    for (uint32_t g = 0; g < G; ++g) {
        const int rc = group_stream(s->ctx, g, &st[g]);
        if (rc != MRT_OK) return rc;
    }
    if (!s->counts_host)
        MRT_HIP(hipHostMalloc(reinterpret_cast<void**>(&s->counts_host), 2ull * kCntWords * 4 * kWfMaxGroups));
    // Events: begin, end, per-group end, per-group checks (two each).
    std::vector<hipEvent_t> evs(2 + 3 * G, nullptr);
    struct Guard {
        std::vector<hipEvent_t>& e;
        ~Guard() {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } guard{evs};
    for (hipEvent_t& e : evs) MRT_HIP(hipEventCreate(&e));
    hipEvent_t b0 = evs[0], b1 = evs[1];
    auto chk = [&](uint32_t g, uint32_t k) { return evs[2 + 3 * g + k]; };
    auto gend = [&](uint32_t g) { return evs[2 + 3 * g + 2]; };
    // Everything queued before the frame on the context stream (uploads,
    // counter clears) precedes every group.
    MRT_HIP(hipEventRecord(b0, st[0]));
    std::vector<WfLoop> L(G);
    for (uint32_t g = 0; g < G; ++g) {
        if (g > 0) MRT_HIP(hipStreamWaitEvent(st[g], b0, 0));
        MRT_HIP(hipMemsetAsync(wbs[g].counts, 0, kCntWords * sizeof(uint32_t), st[g]));
        MRT_HIP(launch_wavefront_start(p, d->first_sample, wbs[g], &L[g], st[g]));
    }
    const uint64_t covered = static_cast<uint64_t>(rendered_tiles(p)) * p.bsx * p.bsy;
    const uint32_t lanes = wbs[0].lanes;
    const uint32_t samples = (d->samples_pixel - d->first_sample + lanes - 1u) / lanes;  // per slot
    // A sample takes at most depth_max + 1 iterations (one per vertex, one to
    // finish); more than that means the queues never drained.
    const uint64_t maxIters = static_cast<uint64_t>(samples) * (static_cast<uint64_t>(d->ray_depth_max) + 2u) +
                              4u * kBatch;
    uint32_t pend[kWfMaxGroups][2][3] = {};
    uint64_t doneCount[kWfMaxGroups] = {};
    bool done[kWfMaxGroups] = {};
    uint32_t checks = 0;
    bool stopReq = false;
    size_t ev = 0;
    // Graph loop: each slot group's batch (kBatch iterations) is captured
    // once on the group's own stream and replayed there until the group
    // drains, so the groups overlap as launched iterations do.  An even
    // number of iterations returns every queue and list to its starting
    // role, so each replay repeats the captured launches exactly; a stop
    // request recaptures with the stop flag set.  The counter snapshots and
    // check events stay outside.  A later frame whose launch arguments are
    // all the same (same scene, descriptor, buffers and launch sizes: a
    // repeated or progressive view) replays the scene's kept graphs without
    // capturing them again.  Each graph is launched on its group's stream
    // behind that group's counter clear and start kernel.
    const bool useGraph = s->ctx->tuning.graph_loop == 1u;
    bool graphStop = false;
    uint64_t graphIters = 0;
    auto graph_key = [&](bool stopFlag) {
        std::vector<uint8_t> k;
        auto put = [&](const void* v, size_t n) {
            const uint8_t* b = static_cast<const uint8_t*>(v);
            k.insert(k.end(), b, b + n);
        };
        const uint32_t flags[4] = {persistent ? 1u : 0u, d->instrument, d->samples_pixel, stopFlag ? 1u : 0u};
        put(flags, sizeof(flags));
        put(&s->dev, sizeof(s->dev));
        put(&p, sizeof(p));
        const void* ptrs[4] = {dbm, dacc, s->counters.p, s->err.p};
        put(ptrs, sizeof(ptrs));
        for (uint32_t g = 0; g < G; ++g) {
            put(&wbs[g], sizeof(WfBuffers));
            put(&scs[g], sizeof(StackCfg));
            put(&st[g], sizeof(hipStream_t));
            const uint32_t blocks[2] = {sized_blocks(trace_blocks, wbs[g].slot_count),
                                        sized_trace_blocks(persist_blocks, wbs[g].slot_count)};
            put(blocks, sizeof(blocks));
        }
        return k;
    };
    auto capture = [&]() -> int {
        std::vector<uint8_t> key = graph_key(stopReq);
        if (s->wf_exec[0] && key == s->wf_key) {
            // Same arguments: replay the kept batches (the loop state
            // advances as the captured launches would have advanced it).
            for (uint32_t k = 0; k < kBatch; ++k)
                for (uint32_t g = 0; g < G; ++g) wavefront_loop_advance(&L[g], wbs[g].sort != 0u);
            graphStop = stopReq;
            return MRT_OK;
        }
        s->wf_key.clear();
        for (uint32_t g = 0; g < kWfMaxGroups; ++g) {
            if (s->wf_exec[g]) MRT_HIP(hipGraphExecDestroy(s->wf_exec[g]));
            s->wf_exec[g] = nullptr;
            if (s->wf_graph[g]) MRT_HIP(hipGraphDestroy(s->wf_graph[g]));
            s->wf_graph[g] = nullptr;
        }
        for (uint32_t g = 0; g < G; ++g) {
            MRT_HIP(hipStreamBeginCapture(st[g], hipStreamCaptureModeThreadLocal));
            hipError_t e = hipSuccess;
            for (uint32_t k = 0; k < kBatch && e == hipSuccess; ++k)
                e = launch_wavefront_iteration(s->dev, s->generic, persistent, d->instrument != 0u, p,
                                               d->samples_pixel, stopReq, dbm, dacc, wbs[g], &L[g],
                                               sized_blocks(trace_blocks, wbs[g].slot_count),
                                               sized_trace_blocks(persist_blocks, wbs[g].slot_count), scs[g],
                                               s->counters.as<unsigned long long>(), s->err.as<uint32_t>(),
                                               nullptr, nullptr, st[g]);
            const hipError_t ec = hipStreamEndCapture(st[g], &s->wf_graph[g]);
            MRT_HIP(e);
            MRT_HIP(ec);
            MRT_HIP(hipGraphInstantiate(&s->wf_exec[g], s->wf_graph[g], nullptr, nullptr, 0));
        }
        s->wf_key = std::move(key);
        graphStop = stopReq;
        return MRT_OK;
    };
    bool graphReady = false;
    while (true) {
        if (!stopReq && s->stop.load()) stopReq = true;
        if (useGraph) {
            if (!graphReady || graphStop != stopReq) {
                const int rc = capture();
                if (rc != MRT_OK) return rc;
                graphReady = true;
            } else {
                for (uint32_t k = 0; k < kBatch; ++k)  // the replay repeats the captured batch
                    for (uint32_t g = 0; g < G; ++g)
                        if (!done[g]) wavefront_loop_advance(&L[g], wbs[g].sort != 0u);
            }
            for (uint32_t g = 0; g < G; ++g) {
                if (done[g]) continue;
                MRT_HIP(hipGraphLaunch(s->wf_exec[g], st[g]));
                graphIters += kBatch;
            }
        }
        for (uint32_t k = 0; k < kBatch && !useGraph; ++k) {
            for (uint32_t g = 0; g < G; ++g) {
                if (done[g]) continue;
                while (ev + 2 > s->trace_events.size()) {
                    hipEvent_t e;
                    MRT_HIP(hipEventCreate(&e));
                    s->trace_events.push_back(e);
                }
                MRT_HIP(launch_wavefront_iteration(s->dev, s->generic, persistent, d->instrument != 0u, p,
                                                   d->samples_pixel, stopReq, dbm, dacc, wbs[g], &L[g],
                                                   sized_blocks(trace_blocks, wbs[g].slot_count),
                                                   sized_trace_blocks(persist_blocks, wbs[g].slot_count), scs[g],
                                                   s->counters.as<unsigned long long>(),
                                                   s->err.as<uint32_t>(), s->trace_events[ev],
                                                   s->trace_events[ev + 1], st[g]));
                ev += 2;
            }
        }
        const uint32_t slot = checks & 1u;
        for (uint32_t g = 0; g < G; ++g) {
            if (done[g]) continue;
            MRT_HIP(hipMemcpyAsync(s->counts_host + (2 * g + slot) * kCntWords, wbs[g].counts,
                                   kCntWords * sizeof(uint32_t), hipMemcpyDeviceToHost, st[g]));
            MRT_HIP(hipEventRecord(chk(g, slot), st[g]));
            std::copy(L[g].pending, L[g].pending + 3, pend[g][slot]);
        }
        if (checks > 0) {
            const uint32_t prev = slot ^ 1u;
            bool all = true;
            for (uint32_t g = 0; g < G; ++g) {
                if (done[g]) continue;
                MRT_HIP(hipEventSynchronize(chk(g, prev)));
                const uint32_t* c = s->counts_host + (2 * g + prev) * kCntWords;
                doneCount[g] = c[kCntDone];
                done[g] = c[pend[g][prev][0]] == 0u && c[pend[g][prev][1]] == 0u && c[pend[g][prev][2]] == 0u;
                all = all && done[g];
            }
            uint64_t finished = 0;
            for (uint32_t g = 0; g < G; ++g) finished += doneCount[g];
            s->progress.store(d->first_sample + static_cast<uint32_t>(finished / covered));
            if (all) break;
        }
        ++checks;
        for (uint32_t g = 0; g < G; ++g)
            if (L[g].iter > maxIters) return fail(MRT_E_HIP, "wavefront queues did not drain");
    }
    for (uint32_t g = 1; g < G; ++g) {
        MRT_HIP(hipEventRecord(gend(g), st[g]));
        MRT_HIP(hipStreamWaitEvent(st[0], gend(g), 0));
    }
    uint32_t* doneMin = wbs[0].counts + (kCntWords - 1u);
    if (wbs[0].rad) {
        MRT_HIP(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(doneMin), static_cast<int>(d->samples_pixel), 1,
                                  st[0]));
        MRT_HIP(launch_wavefront_resolve(p, d->samples_pixel, wbs[0], dbm, dacc, doneMin, st[0]));
    }
    if (wbs[0].aov_s) MRT_HIP(launch_wavefront_aov_resolve(p, d->samples_pixel, wbs[0], daov, st[0]));
    MRT_HIP(hipEventRecord(b1, st[0]));
    MRT_HIP(hipEventSynchronize(b1));
    uint32_t progress = 0;
    if (wbs[0].rad) {
        MRT_HIP(hipMemcpyAsync(&progress, doneMin, sizeof(progress), hipMemcpyDeviceToHost, st[0]));
        MRT_HIP(hipStreamSynchronize(st[0]));
    } else {
        uint64_t finished = 0;
        for (uint32_t g = 0; g < G; ++g) {
            uint32_t c = 0;
            MRT_HIP(hipMemcpyAsync(&c, wbs[g].counts + kCntDone, sizeof(c), hipMemcpyDeviceToHost, st[0]));
            MRT_HIP(hipStreamSynchronize(st[0]));
            finished += c;
        }
        progress = d->first_sample + static_cast<uint32_t>(finished / covered);
    }
    s->progress.store(progress);
    *stopped = stopReq && progress < d->samples_pixel;
    MRT_HIP(hipEventElapsedTime(kernel_ms, b0, b1));
    // Summed launch durations, and the union of the launch intervals.
    float total = 0.0f;
    std::vector<std::pair<float, float>> spans;
    for (size_t i = 0; i < ev; i += 2) {
        float a = 0.0f, b = 0.0f;
        MRT_HIP(hipEventElapsedTime(&a, b0, s->trace_events[i]));
        MRT_HIP(hipEventElapsedTime(&b, b0, s->trace_events[i + 1]));
        total += b - a;
        spans.emplace_back(a, b);
    }
    std::sort(spans.begin(), spans.end());
    float busy = 0.0f, end = -1.0f;
    for (const auto& sp : spans) {
        if (sp.first > end) {
            busy += sp.second - sp.first;
            end = sp.second;
        } else if (sp.second > end) {
            busy += sp.second - end;
            end = sp.second;
        }
    }
    *trace_ms = total;
    *busy_ms = busy;
    *launches = useGraph ? static_cast<uint32_t>(graphIters) : static_cast<uint32_t>(ev / 2);
    return MRT_OK;
}

}  // namespace

namespace mrt {

// Stream of slot group g: the context's stream, then its auxiliary streams
// (created on first use, default priority).
int group_stream(mrt_ctx* c, uint32_t g, hipStream_t* out) {
    if (g == 0) {
        *out = c->stream;
        return MRT_OK;
    }
    hipStream_t& aux = c->aux[g - 1];
    if (!aux) MRT_HIP(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
    *out = aux;
    return MRT_OK;
}

int render_setup(const mrt_scene* s, const mrt_render_desc* d, uint32_t stride_bytes, RenderParams* out,
                 std::vector<uint32_t>* rank_out) {
    int rc;
    if ((rc = check_desc(s, d)) != MRT_OK) return rc;
    const uint32_t w = d->width, h = d->height;
    // Renderer ctor tile sizes, Renderer.cpp:12-29.
    if (w < 16 || h < 16) return fail(MRT_E_INVALID, "width and height must be at least 16");
    if (w % 16u != 0u)
        return fail(MRT_E_UNSUPPORTED, "width must be a multiple of 16 (the Linux front end rounds it, "
                                       "GTK_Interface.cpp:22-30)");
    const uint32_t bsx = w / 16u, bsy = h / 16u;
    if ((w / bsx) * (h / bsy) != kBlocks)
        return fail(MRT_E_UNSUPPORTED, "tile domain (w/bsx)*(h/bsy) must be 256");
    const uint32_t realWidth = stride_bytes / 4u;
    if (stride_bytes % 4u != 0u || realWidth < w || realWidth % w != 0u)
        return fail(MRT_E_UNSUPPORTED, "stride must be a multiple of width * 4 bytes");
    // Tile claim order from the camera's block table (Renderer.cpp:79-88):
    // roundf(block * 256) must visit every tile once.
    std::vector<uint32_t> rank(kBlocks, 0xFFFFFFFFu);
    for (uint32_t t = 0; t < kBlocks; ++t) {
        const float rb = std::roundf(d->tile_values[t] * static_cast<float>(kBlocks));
        if (!(rb >= 0.0f && rb < 256.0f)) return fail(MRT_E_UNSUPPORTED, "tile value out of range");
        const uint32_t r = static_cast<uint32_t>(rb);
        if (rank[r] != 0xFFFFFFFFu) return fail(MRT_E_UNSUPPORTED, "tile values do not form a permutation");
        rank[r] = t;
    }
    const bool compat = d->sampler_pixel.kind == MRT_SAMPLER_HALTON_COMPAT ||
                        d->sampler_pixel.kind == MRT_SAMPLER_STRATIFIED_COMPAT;
    if (compat && d->first_sample != 0) return fail(MRT_E_UNSUPPORTED, "compat samplers render whole frames only");
    RenderParams p = params_from(*d);
    p.bsx = bsx;
    p.bsy = bsy;
    p.real_width = realWidth;
    p.inv_w = 1.0f / static_cast<float>(w);
    p.inv_h = 1.0f / static_cast<float>(h);
    p.pix_w = 0.5f / static_cast<float>(w);
    p.pix_h = 0.5f / static_cast<float>(h);
    p.compat = compat ? 1u : 0u;
    p.pixels_per_pass = kBlocks * bsx * bsy;
    p.tile_rank = nullptr;
    if (rendered_tiles(p) == 0u) return fail(MRT_E_INVALID, "tile_skip skips every tile");
    *out = p;
    *rank_out = std::move(rank);
    return MRT_OK;
}

bool wavefront_eligible(const mrt_scene* s, const mrt_render_desc* d) {
    return d->shader == MRT_SHADER_PATHTRACER && d->samples_light == 1 && !s->dev.has_branches &&
           d->ray_depth_max >= 1 && d->samples_pixel <= 65535u &&
           static_cast<uint64_t>(d->width) * d->height * d->samples_pixel < (1ull << 32);
}

}  // namespace mrt

// The grid of category c built on the GPU into the scene's cell buffers
// (identical to build_grid); *done stays false when the GPU build declines.
// Callers hold ctx->mu.
int build_grid_gpu(mrt_scene* s, int c, const Box& bounds, int32_t gs, GridOut* g, hipStream_t st, bool* done) {
    *done = false;
    mrt_ctx* ctx = s->ctx;
    int rc;
    TriLayout L = kTriPacked;
    uint32_t n = 0;
    if (c == MRT_CAT_TRIANGLE) {
        n = static_cast<uint32_t>(s->tris.size());
        std::vector<float> packed(static_cast<size_t>(n) * 9);
        for (uint32_t i = 0; i < n; ++i) {
            std::memcpy(&packed[9 * static_cast<size_t>(i)], s->tris[i].a, 12);
            std::memcpy(&packed[9 * static_cast<size_t>(i) + 3], s->tris[i].ab, 12);
            std::memcpy(&packed[9 * static_cast<size_t>(i) + 6], s->tris[i].ac, 12);
        }
        if ((rc = ctx->grid_prims.upload(packed.data(), packed.size() * 4, st)) != MRT_OK) return rc;
        MRT_HIP(hipStreamSynchronize(st));  // `packed` is released on return
    } else if (c == MRT_CAT_SPHERE) {
        n = static_cast<uint32_t>(s->sphs.size());
        if ((rc = ctx->grid_prims.upload(s->sphs.data(), s->sphs.size() * sizeof(mrt_sphere), st)) != MRT_OK) return rc;
    } else {
        n = static_cast<uint32_t>(s->plas.size());
        if ((rc = ctx->grid_prims.upload(s->plas.data(), s->plas.size() * sizeof(mrt_plane), st)) != MRT_OK) return rc;
    }
    GridGpuOut o{};
    const hipError_t e =
        gpu_build_grid(c, ctx->grid_prims.as<float>(), L, n, bounds, gs, &ctx->grid_scratch, &o, st);
    if (e == hipErrorNotSupported) return MRT_OK;
    MRT_HIP(e);
    const size_t ncell = static_cast<size_t>(gs) * gs * gs;
    const uint64_t need = buffers_bytes(s) + (ncell + 1 + o.n_refs) * 4ull;
    if (s->ctx->budget && need > s->ctx->budget) return fail(MRT_E_BUDGET, "grid exceeds the HBM budget");
    if ((rc = s->cell_off[c].alloc((ncell + 1) * 4)) != MRT_OK) return rc;
    if ((rc = s->cell_ref[c].alloc(std::max<size_t>(1, o.n_refs) * 4)) != MRT_OK) return rc;
    MRT_HIP(hipMemcpyAsync(s->cell_off[c].p, o.offsets, (ncell + 1) * 4, hipMemcpyDeviceToDevice, st));
    if (o.n_refs) MRT_HIP(hipMemcpyAsync(s->cell_ref[c].p, o.refs, o.n_refs * 4ull, hipMemcpyDeviceToDevice, st));
    grid_meta(bounds, gs, g);
    *done = true;
    return MRT_OK;
}

extern "C" {

MRT_API const char* mrt_last_error(void) { return g_error.c_str(); }

MRT_API int mrt_device_count(int* count) {
    if (!count) return MRT_E_INVALID;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return MRT_OK;
}

MRT_API int mrt_ctx_create(int device, mrt_ctx** out) {
    if (!out) return fail(MRT_E_INVALID, "null output");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(MRT_E_NODEVICE, "no HIP device");
    if (device < 0 || device >= n) return fail(MRT_E_INVALID, "device index out of range");
    MRT_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    MRT_HIP(hipGetDeviceProperties(&prop, device));
    uint32_t* scratch = nullptr;
    MRT_HIP(noalloc_scratch(device, &scratch));
    auto* c = new mrt_ctx;
    c->device = device;
    c->scratch = scratch;
    c->num_cu = static_cast<uint32_t>(prop.multiProcessorCount);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(MRT_E_HIP, "hipStreamCreate failed");
    }
    *out = c;
    return MRT_OK;
}

MRT_API int mrt_ctx_destroy(mrt_ctx* ctx) {
    if (!ctx) return MRT_OK;
    if (ctx->children.load() != 0)
        return fail(MRT_E_INVALID, "context still has scenes or render servers: destroy them first");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamDestroy(ctx->stream);
    for (hipStream_t a : ctx->aux) {
        if (!a) continue;
        (void)hipStreamSynchronize(a);
        (void)hipStreamDestroy(a);
    }
    delete ctx;
    return MRT_OK;
}

MRT_API int mrt_bvh_build_gpu(mrt_ctx* ctx, const float* boxes6, uint32_t n, uint32_t leaf_size, mrt_bvh_node* nodes,
                              uint32_t capacity, uint32_t* count, int mem) {
    if (!ctx || (n && !boxes6) || !count) return fail(MRT_E_INVALID, "null argument");
    if (leaf_size == 0 || leaf_size > 16) return fail(MRT_E_INVALID, "leaf size must be 1..16");
    std::lock_guard<std::mutex> lock(ctx->mu);
    MRT_HIP(hipSetDevice(ctx->device));
    const hipStream_t st = ctx->stream;
    int rc;
    const float* dbox = boxes6;
    if (mem != MRT_MEM_DEVICE && n) {
        if ((rc = ctx->bvh_boxes.upload(boxes6, static_cast<size_t>(n) * 24, st)) != MRT_OK) return rc;
        dbox = ctx->bvh_boxes.as<float>();
    }
    const size_t maxNodes = n ? 2 * static_cast<size_t>(n) - 1 : 1;
    if ((rc = ctx->bvh_nodes.alloc(maxNodes * sizeof(RefNode))) != MRT_OK) return rc;
    uint32_t cnt = 0;
    const hipError_t e = gpu_build_bvh(dbox, n, leaf_size, ctx->bvh_nodes.as<RefNode>(), &cnt, &ctx->bvh_scratch, st);
    if (e == hipErrorNotSupported)
        return fail(MRT_E_UNSUPPORTED, "GPU BVH build: NaN box coordinate or more than 1024 levels (use mrt_bvh_build)");
    MRT_HIP(e);
    *count = cnt;
    static_assert(sizeof(RefNode) == sizeof(mrt_bvh_node), "node layouts differ");
    const size_t copy = static_cast<size_t>(std::min(cnt, capacity)) * sizeof(RefNode);
    if (nodes && copy) {
        MRT_HIP(hipMemcpyAsync(nodes, ctx->bvh_nodes.p, copy,
                               mem == MRT_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st));
        MRT_HIP(hipStreamSynchronize(st));
    }
    return MRT_OK;
}

MRT_API int mrt_grid_build_gpu(mrt_ctx* ctx, int category, const void* prims, uint32_t n, const float bounds6[6],
                               int32_t grid_size, uint32_t* offsets, uint32_t* refs, uint32_t refs_capacity,
                               uint32_t* n_refs, int mem) {
    if (!ctx || (n && !prims) || !bounds6 || !offsets || !n_refs) return fail(MRT_E_INVALID, "null argument");
    if (category < MRT_CAT_PLANE || category > MRT_CAT_TRIANGLE) return fail(MRT_E_INVALID, "unknown category");
    if (grid_size <= 0 || grid_size > 1024) return fail(MRT_E_INVALID, "grid size must be 1..1024");
    std::lock_guard<std::mutex> lock(ctx->mu);
    MRT_HIP(hipSetDevice(ctx->device));
    const hipStream_t st = ctx->stream;
    const size_t rec = category == MRT_CAT_TRIANGLE ? sizeof(mrt_triangle)
                       : category == MRT_CAT_SPHERE ? sizeof(mrt_sphere)
                                                    : sizeof(mrt_plane);
    int rc;
    const void* dp = prims;
    if (mem != MRT_MEM_DEVICE && n) {
        if ((rc = ctx->grid_prims.upload(prims, rec * n, st)) != MRT_OK) return rc;
        dp = ctx->grid_prims.p;
    }
    const Box bounds{ld3(bounds6), ld3(bounds6 + 3)};
    GridGpuOut o{};
    const hipError_t e = gpu_build_grid(category, static_cast<const float*>(dp), kTriAbi, n, bounds, grid_size,
                                        &ctx->grid_scratch, &o, st);
    if (e == hipErrorNotSupported) return fail(MRT_E_UNSUPPORTED, "GPU grid build: non-finite primitive box");
    MRT_HIP(e);
    *n_refs = o.n_refs;
    const hipMemcpyKind kind = mem == MRT_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    const size_t ncell = static_cast<size_t>(grid_size) * grid_size * grid_size;
    MRT_HIP(hipMemcpyAsync(offsets, o.offsets, (ncell + 1) * sizeof(uint32_t), kind, st));
    const size_t nr = std::min(o.n_refs, refs_capacity);
    if (refs && nr) MRT_HIP(hipMemcpyAsync(refs, o.refs, nr * sizeof(uint32_t), kind, st));
    MRT_HIP(hipStreamSynchronize(st));
    return MRT_OK;
}

MRT_API int mrt_ctx_set_budget(mrt_ctx* ctx, uint64_t bytes) {
    if (!ctx) return fail(MRT_E_INVALID, "null context");
    ctx->budget = bytes;
    return MRT_OK;
}

MRT_API int mrt_ctx_set_tuning(mrt_ctx* ctx, const mrt_tuning* t) {
    if (!ctx || !t) return fail(MRT_E_INVALID, "null argument");
    if (t->refill_lanes > 64u) return fail(MRT_E_INVALID, "refill_lanes must be 0-64");
    if (t->fetch_chunk && (t->fetch_chunk < 64u || t->fetch_chunk > 4096u))
        return fail(MRT_E_INVALID, "fetch_chunk must be 0 or 64-4096");
    if (t->launch_share_pct > 100u) return fail(MRT_E_INVALID, "launch_share_pct must be 0-100");
    if (t->max_slot_lanes && (t->max_slot_lanes > 16u || (t->max_slot_lanes & (t->max_slot_lanes - 1u))))
        return fail(MRT_E_INVALID, "max_slot_lanes must be 0 or a power of two up to 16");
    if (t->max_groups > kWfMaxGroups) return fail(MRT_E_INVALID, "max_groups must be 0-3");
    if (t->overlap_loads > 2u) return fail(MRT_E_INVALID, "overlap_loads must be 0-2");
    if (t->graph_loop > 2u) return fail(MRT_E_INVALID, "graph_loop must be 0-2");
    if (t->service_lanes > 64u) return fail(MRT_E_INVALID, "service_lanes must be 0-64");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->tuning = *t;
    return MRT_OK;
}

MRT_API int mrt_ctx_get_tuning(const mrt_ctx* ctx, mrt_tuning* t) {
    if (!ctx || !t) return fail(MRT_E_INVALID, "null argument");
    *t = ctx->tuning;
    return MRT_OK;
}

MRT_API int mrt_ctx_synchronize(mrt_ctx* ctx) {
    if (!ctx) return fail(MRT_E_INVALID, "null context");
    MRT_HIP(hipSetDevice(ctx->device));
    MRT_HIP(hipStreamSynchronize(ctx->stream));
    return MRT_OK;
}

MRT_API int mrt_scene_create(mrt_ctx* ctx, const mrt_scene_desc* d, mrt_scene** out) {
    if (!ctx || !d || !out) return fail(MRT_E_INVALID, "null argument");
    *out = nullptr;
    if ((d->n_triangles && !d->triangles) || (d->n_spheres && !d->spheres) || (d->n_planes && !d->planes) ||
        (d->n_materials && !d->materials) || (d->n_lights && !d->lights))
        return fail(MRT_E_INVALID, "null array with a non-zero count");
    if (d->n_triangles >= (1u << 30) || d->n_spheres >= (1u << 30) || d->n_planes >= (1u << 30) ||
        d->n_lights >= (1u << 30))
        return fail(MRT_E_INVALID, "more than 2^30 primitives in one category");
    for (uint32_t i = 0; i < d->n_triangles; ++i)
        if (d->triangles[i].material_id >= static_cast<int32_t>(d->n_materials))
            return fail(MRT_E_INVALID, "triangle material index out of range");
    for (uint32_t i = 0; i < d->n_spheres; ++i)
        if (d->spheres[i].material_id >= static_cast<int32_t>(d->n_materials))
            return fail(MRT_E_INVALID, "sphere material index out of range");
    for (uint32_t i = 0; i < d->n_planes; ++i)
        if (d->planes[i].material_id >= static_cast<int32_t>(d->n_materials))
            return fail(MRT_E_INVALID, "plane material index out of range");
    const uint64_t estimate = static_cast<uint64_t>(d->n_triangles) * (48 + 80) + d->n_spheres * 20ull +
                              d->n_planes * 32ull + (d->n_materials + d->n_lights) * 64ull + d->n_lights * 160ull;
    if (ctx->budget && estimate > ctx->budget)
        return fail(MRT_E_BUDGET, "scene needs " + std::to_string(estimate) + " bytes, budget " +
                                      std::to_string(ctx->budget));
    std::lock_guard<std::mutex> lock(ctx->mu);
    auto s = std::make_unique<mrt_scene>();
    s->ctx = ctx;
    s->tris.assign(d->triangles, d->triangles + d->n_triangles);
    s->sphs.assign(d->spheres, d->spheres + d->n_spheres);
    s->plas.assign(d->planes, d->planes + d->n_planes);
    s->mats.assign(d->materials, d->materials + d->n_materials);
    s->lights.assign(d->lights, d->lights + d->n_lights);
    s->n_scene_mats = d->n_materials;
    for (const mrt_light& l : s->lights) s->mats.push_back(l.radiance);
    MRT_HIP(hipSetDevice(ctx->device));
    ++ctx->children;  // the scene's destructor gives it back
    s->counted = true;
    const hipStream_t st = ctx->stream;
    int rc;
    // Scene triangles, packed on the device from one upload of their records
    // (below, once the light slots are known), then one slot per area light
    // in light-tree order.
    const size_t nTri = s->tris.size();
    DevBuf raw;
    if ((rc = raw.upload(s->tris.data(), nTri * sizeof(mrt_triangle), st)) != MRT_OK) return rc;
    {
        std::vector<float4> sp(s->sphs.size());
        std::vector<int32_t> sm(s->sphs.size());
        for (size_t i = 0; i < s->sphs.size(); ++i) {
            sp[i] = make_float4(s->sphs[i].center[0], s->sphs[i].center[1], s->sphs[i].center[2],
                                s->sphs[i].sq_radius);
            sm[i] = s->sphs[i].material_id;
        }
        std::vector<float4> pl(2 * s->plas.size());
        for (size_t i = 0; i < s->plas.size(); ++i) {
            float mid;
            std::memcpy(&mid, &s->plas[i].material_id, 4);
            pl[2 * i] = make_float4(s->plas[i].normal[0], s->plas[i].normal[1], s->plas[i].normal[2], mid);
            pl[2 * i + 1] = make_float4(s->plas[i].point[0], s->plas[i].point[1], s->plas[i].point[2], 0.0f);
        }
        std::vector<float4> mt(4 * s->mats.size());
        for (size_t i = 0; i < s->mats.size(); ++i) {
            const mrt_material& m = s->mats[i];
            mt[4 * i + 0] = make_float4(m.le[0], m.le[1], m.le[2], m.kd[0]);
            mt[4 * i + 1] = make_float4(m.kd[1], m.kd[2], m.ks[0], m.ks[1]);
            mt[4 * i + 2] = make_float4(m.ks[2], m.kt[0], m.kt[1], m.kt[2]);
            mt[4 * i + 3] = make_float4(m.ior, 0.0f, 0.0f, 0.0f);
            if (any_gt0(ld3(m.ks)) || any_gt0(ld3(m.kt))) s->dev.has_branches = 1u;
        }
        std::vector<float4> lg(3 * s->lights.size()), ls(5 * s->lights.size());
        std::vector<LightInfo> li(s->lights.size());
        uint32_t nArea = 0;
        std::vector<Box> lboxes;
        std::vector<uint32_t> lidx;
        for (size_t i = 0; i < s->lights.size(); ++i) {
            pack_geo(s->lights[i].triangle, MRT_MAKE_ID(MRT_CAT_LIGHT, static_cast<uint32_t>(i)), &lg[3 * i]);
            pack_shade(s->lights[i].triangle, &ls[5 * i]);
            li[i].kind = s->lights[i].kind;
            li[i].material = s->n_scene_mats + static_cast<uint32_t>(i);
            std::memcpy(li[i].pos, s->lights[i].position, sizeof(li[i].pos));
            li[i].sampler = s->lights[i].sampler;
            if (s->lights[i].kind == MRT_LIGHT_AREA) {
                ++nArea;
                lboxes.push_back(triangle_box(s->lights[i].triangle));
                lidx.push_back(static_cast<uint32_t>(i));
            }
        }
        if ((rc = s->sph.upload(sp.data(), sp.size() * sizeof(float4), st)) != MRT_OK) return rc;
        if ((rc = s->sph_mat.upload(sm.data(), sm.size() * sizeof(int32_t), st)) != MRT_OK) return rc;
        if ((rc = s->pla.upload(pl.data(), pl.size() * sizeof(float4), st)) != MRT_OK) return rc;
        if ((rc = s->mat.upload(mt.data(), mt.size() * sizeof(float4), st)) != MRT_OK) return rc;
        if ((rc = s->light_geo.upload(lg.data(), lg.size() * sizeof(float4), st)) != MRT_OK) return rc;
        if ((rc = s->light_shade.upload(ls.data(), ls.size() * sizeof(float4), st)) != MRT_OK) return rc;
        if ((rc = s->light_info.upload(li.data(), li.size() * sizeof(LightInfo), st)) != MRT_OK) return rc;
        s->dev.n_lights = static_cast<uint32_t>(s->lights.size());
        s->dev.n_area_lights = nArea;
        s->dev.lights_linear = nArea <= kLinearLights ? 1u : 0u;
        // Light slots of the walk tree: light-tree leaf order, or index order
        // for the linear scan.
        std::vector<uint32_t> slotLight = lidx;
        if (!s->dev.lights_linear) {
            BvhOut& lb = s->light_bvh;
            std::vector<uint32_t> perm;
            build_light_bvh(lboxes, &lb, &perm);
            for (uint32_t& p : perm) p = lidx[p];
            if ((rc = upload_bvh(s.get(), 3, lb, st)) != MRT_OK) return rc;
            if ((rc = s->perm[3].upload(perm.data(), perm.size() * sizeof(uint32_t), st)) != MRT_OK) return rc;
            fill_accel_bvh(&s->dev.light_acc, lb, s->nodes[3], &s->perm[3], nArea);
            s->depth_need = std::max(s->depth_need, lb.depth);
            slotLight = perm;
        }
        s->n_light_slots = static_cast<uint32_t>(slotLight.size());
        std::vector<float4> tail(3 * slotLight.size());
        for (size_t k = 0; k < slotLight.size(); ++k)
            std::copy(&lg[3 * slotLight[k]], &lg[3 * slotLight[k]] + 3, &tail[3 * k]);
        if ((rc = s->tri_geo.alloc((3 * nTri + tail.size()) * sizeof(float4))) != MRT_OK) return rc;
        if ((rc = s->tri_shade.alloc(5 * nTri * sizeof(float4))) != MRT_OK) return rc;
        MRT_HIP(gpu_pack_triangles(raw.as<mrt_triangle>(), static_cast<uint32_t>(nTri), s->tri_geo.as<float4>(),
                                   s->tri_shade.as<float4>(), st));
        if (!tail.empty())
            MRT_HIP(hipMemcpyAsync(s->tri_geo.as<float4>() + 3 * nTri, tail.data(), tail.size() * sizeof(float4),
                                   hipMemcpyHostToDevice, st));
        MRT_HIP(hipStreamSynchronize(st));
    }
    s->dev.tri_geo = s->tri_geo.as<float4>();
    s->dev.tri_shade = s->tri_shade.as<float4>();
    s->dev.sph = s->sph.as<float4>();
    s->dev.sph_mat = s->sph_mat.as<int32_t>();
    s->dev.pla = s->pla.as<float4>();
    s->dev.mat = s->mat.as<float4>();
    s->dev.light_geo = s->light_geo.as<float4>();
    s->dev.light_shade = s->light_shade.as<float4>();
    s->dev.light_info = s->light_info.as<LightInfo>();
    s->generic = !s->sphs.empty() || !s->plas.empty();
    s->dev.acc[MRT_CAT_PLANE].count = static_cast<uint32_t>(s->plas.size());
    s->dev.acc[MRT_CAT_SPHERE].count = static_cast<uint32_t>(s->sphs.size());
    s->dev.acc[MRT_CAT_TRIANGLE].count = static_cast<uint32_t>(s->tris.size());
    for (int c = 0; c < 3; ++c) s->dev.acc[c].kind = MRT_NAIVE;
    *out = s.release();
    return MRT_OK;
}

MRT_API int mrt_scene_destroy(mrt_scene* scene) {
    if (!scene) return MRT_OK;
    {
        std::lock_guard<std::mutex> lock(scene->ctx->mu);
        if (scene->servers > 0) return fail(MRT_E_INVALID, "scene is resident in a render server");
    }
    (void)hipSetDevice(scene->ctx->device);
    (void)hipStreamSynchronize(scene->ctx->stream);
    delete scene;
    return MRT_OK;
}

MRT_API int mrt_scene_build(mrt_scene* s, int accelerator, const mrt_accel_params* params) {
    if (!s) return fail(MRT_E_INVALID, "null scene");
    if (accelerator != MRT_NAIVE && accelerator != MRT_REGULAR_GRID && accelerator != MRT_BVH)
        return fail(MRT_E_INVALID, "unknown accelerator");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    MRT_HIP(hipSetDevice(s->ctx->device));
    const hipStream_t st = s->ctx->stream;
    const uint32_t leaf = params && params->leaf_size ? params->leaf_size : 2u;
    if (leaf > 16) return fail(MRT_E_UNSUPPORTED, "leaf size above 16");
    const uint32_t builder = params ? params->builder : static_cast<uint32_t>(MRT_BUILDER_AUTO);
    if (builder > MRT_BUILDER_GPU) return fail(MRT_E_INVALID, "unknown builder");
    const uint32_t counts[3] = {static_cast<uint32_t>(s->plas.size()), static_cast<uint32_t>(s->sphs.size()),
                                static_cast<uint32_t>(s->tris.size())};
    if (s->ctx->budget) {
        // Refused before any device work when the geometry and the smallest
        // accelerator of this kind already exceed the budget: a BVH has at
        // least ceil(n / leaf) - 1 interior nodes (64-byte child pairs, the
        // triangle tree twice with its walk pairs), a grid its cell offsets.
        uint64_t need = s->tri_geo.bytes + s->tri_shade.bytes + s->sph.bytes + s->sph_mat.bytes + s->pla.bytes +
                        s->mat.bytes + s->light_geo.bytes + s->light_shade.bytes + s->light_info.bytes;
        for (int c = 0; c < 3; ++c) {
            if (counts[c] == 0) continue;
            if (accelerator == MRT_BVH)
                need += ((counts[c] + leaf - 1) / leaf - 1) * sizeof(BNode) * (c == MRT_CAT_TRIANGLE ? 2u : 1u);
            else if (accelerator == MRT_REGULAR_GRID)
                need += (static_cast<uint64_t>(grid_size_for(counts[c])) * grid_size_for(counts[c]) *
                             grid_size_for(counts[c]) + 1) * 4u;
        }
        if (need > s->ctx->budget)
            return fail(MRT_E_BUDGET, "accelerator needs at least " + std::to_string(need) + " bytes, budget " +
                                          std::to_string(s->ctx->budget));
    }
    uint32_t lightDepth = s->dev.lights_linear ? 0u : s->dev.light_acc.depth;
    s->depth_need = lightDepth;
    int rc;
    for (int c = 0; c < 3; ++c) {
        Accel& a = s->dev.acc[c];
        const Accel keep_count{};
        a = keep_count;
        a.count = counts[c];
        a.kind = accelerator;
    }
    s->dev.walk_nodes = nullptr;
    s->dev.wide_nodes = nullptr;
    s->dev.cw_nodes = nullptr;
    s->dev.cw_recs = nullptr;
    if (accelerator == MRT_BVH) {
        for (int c = 0; c < 3; ++c) {
            BvhOut b;
            bool onGpu = false;
            if (c == MRT_CAT_TRIANGLE &&
                (builder == MRT_BUILDER_GPU || (builder == MRT_BUILDER_AUTO && counts[c] >= kGpuBuildMin))) {
                bool resident = false;
                if ((rc = build_tri_bvh_resident(s, leaf, st, &b, &onGpu, &resident)) != MRT_OK) return rc;
                if (resident) {
                    s->gpu_built = true;
                    continue;
                }
            }
            if (!onGpu) {
                std::vector<Box> boxes(counts[c]);
                for (uint32_t i = 0; i < counts[c]; ++i)
                    boxes[i] = c == MRT_CAT_TRIANGLE ? triangle_box(s->tris[i])
                               : c == MRT_CAT_SPHERE ? sphere_box(s->sphs[i])
                                                     : plane_box(s->plas[i]);
                build_bvh_reference(boxes, leaf, &b);
            }
            const uint64_t need = buffers_bytes(s) + b.pairs.size() * sizeof(BNode) *
                                                         (c == MRT_CAT_TRIANGLE ? 2u : 1u);
            if (s->ctx->budget && need > s->ctx->budget)
                return fail(MRT_E_BUDGET, "BVH exceeds the HBM budget");
            if ((rc = upload_bvh(s, c, b, st)) != MRT_OK) return rc;
            fill_accel_bvh(&s->dev.acc[c], b, s->nodes[c], nullptr, counts[c]);
            s->depth_need = std::max(s->depth_need, b.depth);
            if (c == MRT_CAT_TRIANGLE && (rc = build_walk_tree(s, b, st, onGpu)) != MRT_OK) return rc;
            if (c == MRT_CAT_TRIANGLE) s->gpu_built = onGpu;
        }
    } else if (accelerator == MRT_REGULAR_GRID) {
        const f3 camMin = params ? ld3(params->camera_min) : f3{0, 0, 0};
        const f3 camMax = params ? ld3(params->camera_max) : f3{0, 0, 0};
        for (int c = 0; c < 3; ++c) {
            auto box_of = [&](uint32_t i) {
                return c == MRT_CAT_TRIANGLE ? triangle_box(s->tris[i])
                       : c == MRT_CAT_SPHERE ? sphere_box(s->sphs[i])
                                             : plane_box(s->plas[i]);
            };
            // Scene::getBounds over the primitives and the camera, widened by
            // Epsilon (Shader.cpp:45-62).
            f3 mn{kRayLengthMax, kRayLengthMax, kRayLengthMax}, mx{-kRayLengthMax, -kRayLengthMax, -kRayLengthMax};
            for (uint32_t i = 0; i < counts[c]; ++i) {
                const Box b = box_of(i);
                mn = vmin(b.mn, mn);
                mx = vmax(b.mx, mx);
            }
            mn = vmin(camMin, mn);
            mx = vmax(camMax, mx);
            const Box bounds{subs(mn, kEpsilon), adds(mx, kEpsilon)};
            const int32_t gs = grid_size_for(counts[c]);
            GridOut g;
            auto hits = [&](uint32_t i, const Box& cell) {
                return c == MRT_CAT_TRIANGLE ? triangle_intersect_box(s->tris[i], cell)
                       : c == MRT_CAT_SPHERE ? sphere_intersect_box(s->sphs[i], cell)
                                             : plane_intersect_box(s->plas[i], cell);
            };
            bool onGpu = false;
            if (counts[c] > 0 &&
                (builder == MRT_BUILDER_GPU || (builder == MRT_BUILDER_AUTO && counts[c] >= kGpuBuildMin))) {
                if ((rc = build_grid_gpu(s, c, bounds, gs, &g, st, &onGpu)) != MRT_OK) return rc;
            }
            if (!onGpu) {
                build_grid(counts[c], bounds, gs, box_of, hits, &g);
                const uint64_t need = buffers_bytes(s) + (g.offsets.size() + g.refs.size()) * 4ull;
                if (s->ctx->budget && need > s->ctx->budget)
                    return fail(MRT_E_BUDGET, "grid exceeds the HBM budget");
                if ((rc = s->cell_off[c].upload(g.offsets.data(), g.offsets.size() * 4, st)) != MRT_OK) return rc;
                if ((rc = s->cell_ref[c].upload(g.refs.data(), g.refs.size() * 4, st)) != MRT_OK) return rc;
            }
            Accel& a = s->dev.acc[c];
            a.cell_off = s->cell_off[c].as<uint32_t>();
            a.cell_ref = s->cell_ref[c].as<uint32_t>();
            a.grid_size = g.grid_size;
            a.grid_shift = g.grid_shift;
            st3(a.ext_mn, g.ext.mn);
            st3(a.sr, g.sr);
            st3(a.cw, g.cw);
        }
    }
    MRT_HIP(hipStreamSynchronize(st));
    s->built = true;
    return MRT_OK;
}

MRT_API int mrt_scene_footprint(const mrt_scene* s, uint64_t* bytes) {
    if (!s || !bytes) return fail(MRT_E_INVALID, "null argument");
    *bytes = buffers_bytes(s);
    return MRT_OK;
}

MRT_API int mrt_trace_closest(mrt_scene* s, const mrt_ray* rays, uint32_t n, mrt_hit* hits, int mem) {
    if (!s || (n && (!rays || !hits))) return fail(MRT_E_INVALID, "null argument");
    if (!s->built) return fail(MRT_E_INVALID, "scene has no accelerator (call mrt_scene_build)");
    if (n == 0) return MRT_OK;
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    MRT_HIP(hipSetDevice(s->ctx->device));
    const hipStream_t st = s->ctx->stream;
    int rc;
    const void* din;
    if ((rc = stage(s->stage_in, rays, n * sizeof(mrt_ray), mem, st, &din)) != MRT_OK) return rc;
    void* dout = hits;
    if (mem != MRT_MEM_DEVICE) {
        if ((rc = s->stage_out.alloc(n * sizeof(mrt_hit))) != MRT_OK) return rc;
        dout = s->stage_out.p;
    }
    if ((rc = prepare_aux(s)) != MRT_OK) return rc;
    const uint32_t blocks = std::min<uint32_t>(cap_blocks(s->ctx), (n + 255u) / 256u);
    StackCfg sc;
    if ((rc = ensure_spill(s, blocks * 256u, &sc)) != MRT_OK) return rc;
    MRT_HIP(launch_trace_closest(s->dev, s->generic, static_cast<const mrt_ray*>(din), n,
                                 static_cast<mrt_hit*>(dout), sc, blocks, false,
                                 s->counters.as<unsigned long long>(), s->err.as<uint32_t>(), st));
    if (mem != MRT_MEM_DEVICE) MRT_HIP(hipMemcpyAsync(hits, dout, n * sizeof(mrt_hit), hipMemcpyDeviceToHost, st));
    return check_err(s);
}

MRT_API int mrt_trace_closest_noalloc(mrt_scene* s, const mrt_ray* rays, uint32_t n, mrt_hit* hits) {
    if (!s || (n && (!rays || !hits))) return fail(MRT_E_INVALID, "null argument");
    if (!s->built) return fail(MRT_E_INVALID, "scene has no accelerator (call mrt_scene_build)");
    if ((reinterpret_cast<uintptr_t>(rays) | reinterpret_cast<uintptr_t>(hits)) & 15u)
        return fail(MRT_E_INVALID, "rays and hits must be 16-byte aligned device memory");
    if (n == 0) return MRT_OK;
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    MRT_HIP(hipSetDevice(s->ctx->device));
    const hipStream_t st = s->ctx->stream;
    uint32_t* scr = s->ctx->scratch;
    if (!scr) return fail(MRT_E_HIP, "no scratch address for this device");
    // The spill of one launch must fit the scratch: as many 256-lane
    // workgroups as it holds, each lane walking rays grid-stride.
    uint32_t blocks = std::min<uint32_t>(cap_blocks(s->ctx), (n + 255u) / 256u);
    StackCfg sc{nullptr, 0u, 0u};
    const uint32_t need = s->depth_need + 1u;
    if (need > kStackLds) {
        sc.cap_spill = need - kStackLds;
        const uint64_t fit = (kNoAllocScratchWords - kNoAllocReserved) / (256ull * sc.cap_spill);
        if (fit == 0)
            return fail(MRT_E_BUDGET, "tree depth " + std::to_string(s->depth_need) + " needs " +
                                          std::to_string(256ull * sc.cap_spill) +
                                          " scratch words per workgroup; the allocation-free scratch holds " +
                                          std::to_string(kNoAllocScratchWords - kNoAllocReserved));
        blocks = static_cast<uint32_t>(std::min<uint64_t>(blocks, fit));
        sc.spill = scr + kNoAllocReserved;
        sc.spill_stride = blocks * 256u;
    }
    uint32_t* err = scr;
    MRT_HIP(launch_fill_u32(err, 1u, 0u, st));
    MRT_HIP(launch_trace_closest(s->dev, s->generic, rays, n, hits, sc, blocks, false, nullptr, err, st));
    uint32_t e = 0;
    MRT_HIP(hipMemcpyAsync(&e, err, sizeof(e), hipMemcpyDeviceToHost, st));
    MRT_HIP(hipStreamSynchronize(st));
    if (e & 1u) return fail(MRT_E_INVALID, "traversal stack overflow (BVH deeper than the stack budget)");
    return MRT_OK;
}

MRT_API int mrt_trace_any(mrt_scene* s, const mrt_ray* rays, uint32_t n, uint8_t* occluded, int mem) {
    if (!s || (n && (!rays || !occluded))) return fail(MRT_E_INVALID, "null argument");
    if (!s->built) return fail(MRT_E_INVALID, "scene has no accelerator (call mrt_scene_build)");
    if (n == 0) return MRT_OK;
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    MRT_HIP(hipSetDevice(s->ctx->device));
    const hipStream_t st = s->ctx->stream;
    int rc;
    const void* din;
    if ((rc = stage(s->stage_in, rays, n * sizeof(mrt_ray), mem, st, &din)) != MRT_OK) return rc;
    void* dout = occluded;
    if (mem != MRT_MEM_DEVICE) {
        if ((rc = s->stage_out.alloc(n)) != MRT_OK) return rc;
        dout = s->stage_out.p;
    }
    if ((rc = prepare_aux(s)) != MRT_OK) return rc;
    const uint32_t blocks = std::min<uint32_t>(cap_blocks(s->ctx), (n + 255u) / 256u);
    StackCfg sc;
    if ((rc = ensure_spill(s, blocks * 256u, &sc)) != MRT_OK) return rc;
    MRT_HIP(launch_trace_any(s->dev, s->generic, static_cast<const mrt_ray*>(din), n, static_cast<uint8_t*>(dout),
                             sc, blocks, s->err.as<uint32_t>(), st));
    if (mem != MRT_MEM_DEVICE) MRT_HIP(hipMemcpyAsync(occluded, dout, n, hipMemcpyDeviceToHost, st));
    return check_err(s);
}

MRT_API int mrt_shade_rays(mrt_scene* s, const mrt_render_desc* d, const mrt_ray* rays, const int32_t* depths,
                           uint32_t n, uint32_t sample, float* rgb, uint8_t* hit_light, int mem) {
    if (!s || (n && (!rays || !rgb || !hit_light))) return fail(MRT_E_INVALID, "null argument");
    if (!s->built) return fail(MRT_E_INVALID, "scene has no accelerator (call mrt_scene_build)");
    int rc;
    if ((rc = check_desc(s, d)) != MRT_OK) return rc;
    if (d->sampler_pixel.kind >= MRT_SAMPLER_HALTON_COMPAT)
        return fail(MRT_E_UNSUPPORTED, "compat pixel samplers apply to mrt_render only");
    if (n == 0) return MRT_OK;
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    MRT_HIP(hipSetDevice(s->ctx->device));
    const hipStream_t st = s->ctx->stream;
    const void *din, *ddep = nullptr;
    if ((rc = stage(s->stage_in, rays, n * sizeof(mrt_ray), mem, st, &din)) != MRT_OK) return rc;
    if (depths && (rc = stage(s->stage_aux, depths, n * sizeof(int32_t), mem, st, &ddep)) != MRT_OK) return rc;
    float* drgb = rgb;
    uint8_t* dflag = hit_light;
    if (mem != MRT_MEM_DEVICE) {
        if ((rc = s->stage_out.alloc(n * (3 * sizeof(float) + 1))) != MRT_OK) return rc;
        drgb = s->stage_out.as<float>();
        dflag = reinterpret_cast<uint8_t*>(drgb + 3ull * n);
    }
    if ((rc = prepare_aux(s)) != MRT_OK) return rc;
    const RenderParams p = params_from(*d);
    const uint32_t blocks = std::min<uint32_t>(cap_blocks(s->ctx), (n + 255u) / 256u);
    StackCfg sc;
    if ((rc = ensure_spill(s, blocks * 256u, &sc)) != MRT_OK) return rc;
    MRT_HIP(launch_shade_rays(s->dev, s->generic, s->dev.has_branches != 0u, p, static_cast<const mrt_ray*>(din),
                              static_cast<const int32_t*>(ddep), n, sample, drgb, dflag, sc, blocks,
                              s->err.as<uint32_t>(), st));
    if (mem != MRT_MEM_DEVICE) {
        MRT_HIP(hipMemcpyAsync(rgb, drgb, n * 3 * sizeof(float), hipMemcpyDeviceToHost, st));
        MRT_HIP(hipMemcpyAsync(hit_light, dflag, n, hipMemcpyDeviceToHost, st));
    }
    return check_err(s);
}

MRT_API int mrt_camera_rays(mrt_ctx* ctx, const mrt_camera* cam, const float* uvdd, uint32_t n, mrt_ray* rays,
                            int mem) {
    if (!ctx || !cam || (n && (!uvdd || !rays))) return fail(MRT_E_INVALID, "null argument");
    if (n == 0) return MRT_OK;
    std::lock_guard<std::mutex> lock(ctx->mu);
    MRT_HIP(hipSetDevice(ctx->device));
    DevBuf in, out;
    const void* din = uvdd;
    void* dout = rays;
    int rc;
    if (mem != MRT_MEM_DEVICE) {
        if ((rc = in.upload(uvdd, n * 4 * sizeof(float), ctx->stream)) != MRT_OK) return rc;
        if ((rc = out.alloc(n * sizeof(mrt_ray))) != MRT_OK) return rc;
        din = in.p;
        dout = out.p;
    }
    MRT_HIP(launch_camera_rays(*cam, static_cast<const float*>(din), n, static_cast<mrt_ray*>(dout), ctx->stream));
    if (mem != MRT_MEM_DEVICE)
        MRT_HIP(hipMemcpyAsync(rays, dout, n * sizeof(mrt_ray), hipMemcpyDeviceToHost, ctx->stream));
    MRT_HIP(hipStreamSynchronize(ctx->stream));
    return MRT_OK;
}

MRT_API int mrt_intersect_pairs(mrt_ctx* ctx, int category, const void* prims, const mrt_ray* rays, uint32_t n,
                                mrt_hit* closest, uint8_t* any, int mem) {
    if (!ctx || (n && (!prims || !rays))) return fail(MRT_E_INVALID, "null argument");
    if (category < MRT_CAT_PLANE || category > MRT_CAT_LIGHT) return fail(MRT_E_INVALID, "category");
    if (n == 0) return MRT_OK;
    const size_t psize = category == MRT_CAT_PLANE    ? sizeof(mrt_plane)
                         : category == MRT_CAT_SPHERE ? sizeof(mrt_sphere)
                                                      : sizeof(mrt_triangle);
    std::lock_guard<std::mutex> lock(ctx->mu);
    MRT_HIP(hipSetDevice(ctx->device));
    DevBuf bp, br, bc, ba;
    const void* dp = prims;
    const void* dr = rays;
    mrt_hit* dc = closest;
    uint8_t* da = any;
    int rc;
    if (mem != MRT_MEM_DEVICE) {
        if ((rc = bp.upload(prims, n * psize, ctx->stream)) != MRT_OK) return rc;
        if ((rc = br.upload(rays, n * sizeof(mrt_ray), ctx->stream)) != MRT_OK) return rc;
        dp = bp.p;
        dr = br.p;
        if (closest) {
            if ((rc = bc.alloc(n * sizeof(mrt_hit))) != MRT_OK) return rc;
            dc = bc.as<mrt_hit>();
        }
        if (any) {
            if ((rc = ba.alloc(n)) != MRT_OK) return rc;
            da = ba.as<uint8_t>();
        }
    }
    MRT_HIP(launch_intersect_pairs(category, dp, static_cast<const mrt_ray*>(dr), n, dc, da, ctx->stream));
    if (mem != MRT_MEM_DEVICE) {
        if (closest) MRT_HIP(hipMemcpyAsync(closest, dc, n * sizeof(mrt_hit), hipMemcpyDeviceToHost, ctx->stream));
        if (any) MRT_HIP(hipMemcpyAsync(any, da, n, hipMemcpyDeviceToHost, ctx->stream));
    }
    MRT_HIP(hipStreamSynchronize(ctx->stream));
    return MRT_OK;
}

MRT_API int mrt_box_rays(mrt_ctx* ctx, const float* boxes6, const mrt_ray* rays, uint32_t n, uint8_t* hit,
                         float* entry, int mem) {
    if (!ctx || (n && (!boxes6 || !rays || !hit))) return fail(MRT_E_INVALID, "null argument");
    if (n == 0) return MRT_OK;
    std::lock_guard<std::mutex> lock(ctx->mu);
    MRT_HIP(hipSetDevice(ctx->device));
    DevBuf bb, br, bh, be;
    const void* db = boxes6;
    const void* dr = rays;
    uint8_t* dh = hit;
    float* de = entry;
    int rc;
    if (mem != MRT_MEM_DEVICE) {
        if ((rc = bb.upload(boxes6, n * 6ull * sizeof(float), ctx->stream)) != MRT_OK) return rc;
        if ((rc = br.upload(rays, n * sizeof(mrt_ray), ctx->stream)) != MRT_OK) return rc;
        if ((rc = bh.alloc(n)) != MRT_OK) return rc;
        db = bb.p;
        dr = br.p;
        dh = bh.as<uint8_t>();
        if (entry) {
            if ((rc = be.alloc(n * sizeof(float))) != MRT_OK) return rc;
            de = be.as<float>();
        }
    }
    MRT_HIP(launch_box_rays(static_cast<const float*>(db), static_cast<const mrt_ray*>(dr), n, dh, de, ctx->stream));
    if (mem != MRT_MEM_DEVICE) {
        MRT_HIP(hipMemcpyAsync(hit, dh, n, hipMemcpyDeviceToHost, ctx->stream));
        if (entry) MRT_HIP(hipMemcpyAsync(entry, de, n * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    }
    MRT_HIP(hipStreamSynchronize(ctx->stream));
    return MRT_OK;
}

// Renderer::renderFrame, and with aov the auxiliary outputs of the same
// frame (recorded by the wavefront pipeline from the camera rays it traces,
// or by the AOV kernel after the render).
static int render_impl(mrt_scene* s, const mrt_render_desc* d, uint32_t* bitmap, uint32_t stride_bytes,
                       float* accum, float* aov, int mem, mrt_render_stats* stats) {
    if (!s || !bitmap) return fail(MRT_E_INVALID, "null argument");
    if (!s->built) return fail(MRT_E_INVALID, "scene has no accelerator (call mrt_scene_build)");
    int rc;
    RenderParams p;
    std::vector<uint32_t> rank;
    if ((rc = render_setup(s, d, stride_bytes, &p, &rank)) != MRT_OK) return rc;
    const uint32_t w = d->width, h = d->height;
    const uint32_t bsx = p.bsx, bsy = p.bsy;
    const uint32_t realWidth = p.real_width;
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    MRT_HIP(hipSetDevice(s->ctx->device));
    const hipStream_t st = s->ctx->stream;
    if ((rc = s->tile_rank.upload(rank.data(), rank.size() * sizeof(uint32_t), st)) != MRT_OK) return rc;
    p.tile_rank = s->tile_rank.as<uint32_t>();
    const size_t bmBytes = static_cast<size_t>(h) * realWidth * sizeof(uint32_t);
    const size_t accBytes = static_cast<size_t>(w) * h * 3 * sizeof(float);
    uint32_t* dbm = bitmap;
    float* dacc = accum;
    if (mem != MRT_MEM_DEVICE) {
        if ((rc = s->stage_in.upload(bitmap, bmBytes, st)) != MRT_OK) return rc;
        dbm = s->stage_in.as<uint32_t>();
        if (accum) {
            if ((rc = s->stage_out.upload(accum, accBytes, st)) != MRT_OK) return rc;
            dacc = s->stage_out.as<float>();
        }
    }
    const size_t aovBytes = static_cast<size_t>(w) * h * 7 * sizeof(float);
    float* daov = aov;
    if (aov) {
        if (mem != MRT_MEM_DEVICE) {
            if ((rc = s->stage_aux.alloc(aovBytes)) != MRT_OK) return rc;
            daov = s->stage_aux.as<float>();
        }
        MRT_HIP(hipMemsetAsync(daov, 0, aovBytes, st));
    }
    if ((rc = prepare_aux(s)) != MRT_OK) return rc;
    // Kernel shape.  AUTO: the persistent megakernel for PathTracer frames
    // of BVH scenes of triangles and lights below 6 x 2^20 pixel samples (one
    // launch, no iterations: on the 1M-triangle soup 3.95 / 9.81 / 17.2 / 29.4
    // ms for 256x256 x 1, 640x480 x 4, 1280x720 x 4 and 640x480 x 16 samples
    // against 7.49 / 13.9 / 20.2 / 34.3 ms for the wavefront), the wavefront
    // above (its trace kernel runs 7 waves per SIMD against the megakernel's
    // 4: 29.4 / 31.9 / 53.9 / 106.3 ms for 1280x720 x 8, 1920x1080 x 4, 8 and
    // 16 samples against 30.4 / 33.2 / 62.3 / 124.8 ms), for the other frames
    // it can run and for frames with auxiliary outputs (it records them while
    // shading), else the megakernel per sample pass.
    const bool wfEligible = wavefront_eligible(s, d);
    const Accel& triAcc = s->dev.acc[MRT_CAT_TRIANGLE];
    const bool megaEligible = wfEligible && d->persistent != 2u && !s->generic && triAcc.kind == MRT_BVH &&
                              s->dev.walk_nodes != nullptr && s->dev.cw_nodes != nullptr;
    bool useWavefront = false;
    if (d->kernel == MRT_KERNEL_WAVEFRONT) {
        if (!wfEligible)
            return fail(MRT_E_UNSUPPORTED, "the wavefront kernels need the PathTracer, samples_light == 1, no "
                                           "reflective or transmissive materials, ray_depth_max >= 1 and at most "
                                           "65535 samples per pixel");
        useWavefront = true;
    } else if (d->kernel == MRT_KERNEL_AUTO) {
        const uint64_t pixelSamples =
            static_cast<uint64_t>(rendered_tiles(p)) * p.bsx * p.bsy *
            (d->samples_pixel > d->first_sample ? d->samples_pixel - d->first_sample : 0u);
        useWavefront = wfEligible && (!megaEligible || aov != nullptr || pixelSamples >= kMegaMaxPixelSamples);
    }
    const uint32_t traceBlocks = cap_blocks(s->ctx) / 2u;
    // Persistent refilling trace: over the combined walk tree for BVH scenes,
    // over a resumable grid walk for RegularGrid scenes (triangles and lights
    // only either way).
    const bool persistent =
        useWavefront && d->persistent != 2u && !s->generic &&
        ((triAcc.kind == MRT_BVH && s->dev.walk_nodes != nullptr && s->dev.cw_nodes != nullptr) ||
         (triAcc.kind == MRT_REGULAR_GRID && triAcc.count > 0 && triAcc.cell_off != nullptr));
    // The persistent megakernel: megakernel frames the wavefront could run,
    // over the combined walk tree.
    const bool megaP = !useWavefront && megaEligible;
    const uint32_t megaBlocks = megaP ? mega_blocks(s->ctx->num_cu ? s->ctx->num_cu : 256u) : 0u;
    const uint32_t blocks = useWavefront ? traceBlocks : megaP ? megaBlocks : render_pass_blocks(p);
    // Wavefront slot groups run concurrently, each with its own spill region;
    // sample-parallel frames give every pixel `lanes` slots.
    const uint32_t pixelSlots = useWavefront || megaP ? wavefront_slots(p) : 0u;
    const mrt_tuning& tune = s->ctx->tuning;
    const uint32_t lanes =
        useWavefront || megaP ? wavefront_lanes(pixelSlots, d->samples_pixel - d->first_sample,
                                                tune.max_slot_lanes ? tune.max_slot_lanes : kDefaultSlotLanes)
                              : 1u;
    const uint32_t slots = pixelSlots * lanes;
    const uint32_t groups = useWavefront ? wavefront_groups(slots, tune.max_groups ? tune.max_groups : kWfMaxGroups)
                                         : 1u;
    const uint32_t persistBlocks =
        persistent ? wavefront_persistent_blocks(s->ctx->num_cu ? s->ctx->num_cu : 256u, traceBlocks, groups,
                                                 tune.launch_share_pct)
                   : 0u;
    StackCfg sc;
    if ((rc = ensure_spill(s, blocks * 256u * groups, &sc)) != MRT_OK) return rc;
    std::vector<StackCfg> scs(groups, sc);
    for (uint32_t g = 0; g < groups; ++g) {
        scs[g].spill_stride = blocks * 256u;
        if (sc.spill) scs[g].spill = sc.spill + static_cast<size_t>(g) * sc.cap_spill * blocks * 256u;
    }
    sc = scs[0];
    std::vector<WfBuffers> wbs;
    if (useWavefront) {
        const size_t scanBytes = wavefront_scan_bytes();
        if ((rc = s->wf_state.alloc(wavefront_state_bytes(slots))) != MRT_OK) return rc;
        float4* state = s->wf_state.as<float4>();
        float* rad = nullptr;
        if (lanes > 1u) {
            const size_t n = 3ull * (d->samples_pixel - d->first_sample) * pixelSlots;
            if ((rc = s->wf_rad.alloc(n * sizeof(float))) != MRT_OK) return rc;
            rad = s->wf_rad.as<float>();
            MRT_HIP(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(rad), static_cast<int>(kRadPending), n, st));
        }
        // Auxiliary outputs: one first-hit record per sample and pixel slot
        // (2 float4) up to 2 GiB; larger frames run the AOV kernel after the
        // render.
        float4* aovS = nullptr;
        const size_t aovSBytes = 32ull * (d->samples_pixel - d->first_sample) * pixelSlots;
        if (aov && aovSBytes <= (2ull << 30)) {
            if ((rc = s->wf_aov.alloc(aovSBytes)) != MRT_OK) return rc;
            aovS = s->wf_aov.as<float4>();
            MRT_HIP(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(aovS), static_cast<int>(kRadPending),
                                      aovSBytes / 4, st));
        }
        // Key box: the triangle BVH's root box (the unit cube without triangles).
        f3 mn{0.0f, 0.0f, 0.0f}, mx{1.0f, 1.0f, 1.0f};
        const Accel& ta = s->dev.acc[MRT_CAT_TRIANGLE];
        if (ta.kind == MRT_BVH && ta.count > 0) {
            mn = ld3(ta.root_mn);
            mx = ld3(ta.root_mx);
        }
        // Groups own whole waves of slots (64 slots, one 8x8 pixel block each).
        const uint32_t waves = slots / 64u;
        for (uint32_t g = 0; g < groups; ++g) {
            WfBuffers wb{};
            const uint32_t w0 = waves * g / groups, w1 = waves * (g + 1) / groups;
            wb.slot_base = 64u * w0;
            wb.slot_count = 64u * (w1 - w0);
            const uint32_t n = wb.slot_count;
            if ((rc = s->wf_group[g].alloc(wavefront_group_bytes(n) + scanBytes + 256)) != MRT_OK) return rc;
            float4* base = s->wf_group[g].as<float4>();
            wb.sA = state;
            wb.sB = state + slots;
            wb.sC = state + 2ull * slots;
            wb.queue_a = base;
            wb.queue_b = base + 2ull * n;
            wb.hits = base + 4ull * n;
            wb.shadow = base + 5ull * n;
            uint32_t* u = reinterpret_cast<uint32_t*>(base + 9ull * n);
            wb.counts = u;
            wb.ended[0] = u + 64;
            wb.ended[1] = wb.ended[0] + n;
            wb.keys = wb.ended[1] + n;
            wb.hist = wb.keys + n;
            wb.offs = wb.hist + kSortBins;
            const uintptr_t t =
                (reinterpret_cast<uintptr_t>(wb.offs + kSortBins) + 255u) & ~static_cast<uintptr_t>(255u);
            wb.scan_temp = reinterpret_cast<void*>(t);
            wb.scan_bytes = scanBytes;
            wb.sort = d->ray_sort == 1u ? 1u : 0u;
            wb.refill_lanes = tune.refill_lanes ? tune.refill_lanes : kDefaultRefillLanes;
            wb.fetch_chunk = tune.fetch_chunk ? tune.fetch_chunk : kDefaultFetchChunk;
            wb.overlap = wavefront_overlap(2ull * wb.slot_count, persistBlocks, tune.overlap_loads) ? 1u : 0u;
            wb.rad = rad;
            wb.aov_s = aovS;
            wb.lanes = lanes;
            wb.pixel_slots = pixelSlots;
            wb.first = d->first_sample;
            for (int k = 0; k < 3; ++k) {
                const float e = comp(mx, k) - comp(mn, k);
                wb.box_mn[k] = comp(mn, k);
                wb.box_inv[k] = e > 0.0f ? 32.0f / e : 0.0f;
            }
            MRT_HIP(hipMemsetAsync(wb.hist, 0, kSortBins * sizeof(uint32_t), st));
            wbs.push_back(wb);
        }
    }
    s->stop.store(false);
    s->progress.store(d->first_sample);
    // Events around the whole render and around every pass launch (the pass
    // kernel is launched on this context's stream, so these events time it).
    hipEvent_t e0, e1, k0, k1;
    MRT_HIP(hipEventCreate(&e0));
    MRT_HIP(hipEventCreate(&e1));
    MRT_HIP(hipEventCreate(&k0));
    MRT_HIP(hipEventCreate(&k1));
    MRT_HIP(hipEventRecord(e0, st));
    bool stopped = false;
    float kernel_ms = 0.0f, trace_ms = 0.0f, trace_busy = 0.0f;
    uint32_t traceLaunches = 0;
    if (useWavefront && d->first_sample < d->samples_pixel) {
        if ((rc = run_wavefront(s, d, p, wbs, persistent, traceBlocks, persistBlocks, scs, dbm, dacc, daov, &stopped,
                                &kernel_ms, &trace_ms, &trace_busy, &traceLaunches)) != MRT_OK)
            return rc;
    }
    if (megaP && d->first_sample < d->samples_pixel && !s->stop.load()) {
        // One launch for the frame; per-sample radiance and the resolve
        // when a pixel's samples run on several lanes.
        if ((rc = s->mega_ctl.alloc(256)) != MRT_OK) return rc;
        uint32_t* ctl = s->mega_ctl.as<uint32_t>();
        MRT_HIP(hipMemsetAsync(ctl, 0, 256, st));
        float* rad = nullptr;
        if (lanes > 1u) {
            const size_t n = 3ull * (d->samples_pixel - d->first_sample) * pixelSlots;
            if ((rc = s->wf_rad.alloc(n * sizeof(float))) != MRT_OK) return rc;
            rad = s->wf_rad.as<float>();
            MRT_HIP(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(rad), static_cast<int>(kRadPending), n, st));
        }
        MegaArgs a{};
        a.p = p;
        a.first = d->first_sample;
        a.spp = d->samples_pixel;
        a.lanes = lanes;
        a.pixel_slots = pixelSlots;
        a.rad = rad;
        a.bitmap = dbm;
        a.accum = dacc;
        a.fetch = ctl;
        a.done = ctl + 16;
        a.service_lanes = tune.service_lanes ? tune.service_lanes : kDefaultServiceLanes;
        a.refill_lanes = tune.refill_lanes ? tune.refill_lanes : kDefaultMegaRefillLanes;
        a.fetch_chunk = tune.fetch_chunk ? tune.fetch_chunk : kDefaultMegaFetchChunk;
        // Path state between walks in device memory rather than registers:
        // the kernel needs about 160 registers either way (3 waves per SIMD),
        // and this way measured 857 against 836 Mrays/s on the 1M-triangle
        // 1080p frame.
        if ((rc = s->mega_state.alloc(7ull * sizeof(float4) * megaBlocks * 256u)) != MRT_OK) return rc;
        a.state = s->mega_state.as<float4>();
        // ms_kernel: the megakernel alone (the resolve follows it).
        MRT_HIP(hipEventRecord(k0, st));
        MRT_HIP(launch_mega_p(s->dev, d->instrument != 0u, a, megaBlocks, sc, s->counters.as<unsigned long long>(),
                              s->err.as<uint32_t>(), st));
        MRT_HIP(hipEventRecord(k1, st));
        if (rad) {
            WfBuffers wb{};
            wb.rad = rad;
            wb.lanes = lanes;
            wb.pixel_slots = pixelSlots;
            wb.first = d->first_sample;
            uint32_t* doneMin = ctl + 32;
            MRT_HIP(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(doneMin), static_cast<int>(d->samples_pixel),
                                      1, st));
            MRT_HIP(launch_wavefront_resolve(p, d->samples_pixel, wb, dbm, dacc, doneMin, st));
        }
        MRT_HIP(hipEventSynchronize(k1));
        MRT_HIP(hipEventElapsedTime(&kernel_ms, k0, k1));
        s->progress.store(d->samples_pixel);
    } else if (megaP && s->stop.load()) {
        stopped = true;
    }
    for (uint32_t sample = d->first_sample; !useWavefront && !megaP && sample < d->samples_pixel; ++sample) {
        if (s->stop.load()) {
            stopped = true;
            break;
        }
        MRT_HIP(hipEventRecord(k0, st));
        MRT_HIP(launch_render_pass(s->dev, s->generic, s->dev.has_branches != 0u, d->instrument != 0u, p, sample, dbm,
                                   dacc, sc, s->counters.as<unsigned long long>(), s->err.as<uint32_t>(), st));
        MRT_HIP(hipEventRecord(k1, st));
        MRT_HIP(hipEventSynchronize(k1));
        float kms = 0.0f;
        MRT_HIP(hipEventElapsedTime(&kms, k0, k1));
        kernel_ms += kms;
        s->progress.store(sample + 1);
    }
    MRT_HIP(hipEventRecord(e1, st));
    MRT_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    MRT_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(k0);
    (void)hipEventDestroy(k1);
    const bool aovDone = useWavefront && !wbs.empty() && wbs[0].aov_s != nullptr;
    if (aov && !aovDone && d->first_sample < d->samples_pixel) {
        StackCfg asc;
        if ((rc = ensure_spill(s, render_pass_blocks(p) * 256u, &asc)) != MRT_OK) return rc;
        MRT_HIP(launch_aov(s->dev, s->generic, p, d->first_sample, d->samples_pixel, daov, asc, s->err.as<uint32_t>(),
                           st));
    }
    if (mem != MRT_MEM_DEVICE) {
        MRT_HIP(hipMemcpyAsync(bitmap, dbm, bmBytes, hipMemcpyDeviceToHost, st));
        if (accum) MRT_HIP(hipMemcpyAsync(accum, dacc, accBytes, hipMemcpyDeviceToHost, st));
        if (aov) MRT_HIP(hipMemcpyAsync(aov, daov, aovBytes, hipMemcpyDeviceToHost, st));
    }
    unsigned long long cnt[16] = {0};
    MRT_HIP(hipMemcpyAsync(cnt, s->counters.p, sizeof(cnt), hipMemcpyDeviceToHost, st));
    if ((rc = check_err(s)) != MRT_OK) return rc;
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->rays_camera = cnt[0];
        stats->rays_extension = cnt[1];
        stats->rays_shadow = cnt[2];
        stats->rays_specular = cnt[3];
        stats->node_visits = cnt[4] + cnt[5] + cnt[6];
        stats->prim_tests = cnt[7] + cnt[8] + cnt[9];
        stats->light_node_visits = cnt[5];
        stats->light_prim_tests = cnt[8];
        stats->shadow_node_visits = cnt[6];
        stats->shadow_prim_tests = cnt[9];
        stats->samples_done = s->progress.load();
        stats->covered_pixels = rendered_tiles(p) * bsx * bsy;
        stats->ms_render = ms;
        stats->ms_kernel = kernel_ms;
        stats->ms_trace = trace_ms;
        stats->trace_launches = traceLaunches;
        stats->trace_lane_iters = cnt[10];
        stats->trace_lane_slots = cnt[11];
        stats->ms_trace_busy = trace_busy;
        stats->trace_groups = useWavefront ? static_cast<uint32_t>(wbs.size()) : 0u;
        stats->slot_lanes = useWavefront || megaP ? lanes : 0u;
        stats->persistent_megakernel = megaP ? 1u : 0u;
    }
    if (stopped) return fail(MRT_E_STOPPED, "render stopped");
    return MRT_OK;
}

MRT_API int mrt_render(mrt_scene* s, const mrt_render_desc* d, uint32_t* bitmap, uint32_t stride_bytes, float* accum,
                       int mem, mrt_render_stats* stats) {
    return render_impl(s, d, bitmap, stride_bytes, accum, nullptr, mem, stats);
}

MRT_API int mrt_render_with_aovs(mrt_scene* s, const mrt_render_desc* d, uint32_t* bitmap, uint32_t stride_bytes,
                                 float* accum, float* aov, int mem, mrt_render_stats* stats) {
    if (!aov) return fail(MRT_E_INVALID, "null aov");
    return render_impl(s, d, bitmap, stride_bytes, accum, aov, mem, stats);
}

MRT_API int mrt_render_aovs(mrt_scene* s, const mrt_render_desc* d, float* aov, int mem) {
    if (!s || !aov) return fail(MRT_E_INVALID, "null argument");
    if (!s->built) return fail(MRT_E_INVALID, "scene has no accelerator (call mrt_scene_build)");
    int rc;
    RenderParams p;
    std::vector<uint32_t> rank;
    if ((rc = render_setup(s, d, d ? d->width * 4u : 0u, &p, &rank)) != MRT_OK) return rc;
    if (d->first_sample >= d->samples_pixel) return fail(MRT_E_INVALID, "first_sample must be below samples_pixel");
    std::lock_guard<std::mutex> lock(s->ctx->mu);
    MRT_HIP(hipSetDevice(s->ctx->device));
    const hipStream_t st = s->ctx->stream;
    if ((rc = s->tile_rank.upload(rank.data(), rank.size() * sizeof(uint32_t), st)) != MRT_OK) return rc;
    p.tile_rank = s->tile_rank.as<uint32_t>();
    const size_t bytes = static_cast<size_t>(d->width) * d->height * 7 * sizeof(float);
    float* dout = aov;
    if (mem != MRT_MEM_DEVICE) {
        if ((rc = s->stage_out.alloc(bytes)) != MRT_OK) return rc;
        dout = s->stage_out.as<float>();
    }
    MRT_HIP(hipMemsetAsync(dout, 0, bytes, st));
    if ((rc = prepare_aux(s)) != MRT_OK) return rc;
    StackCfg sc;
    if ((rc = ensure_spill(s, render_pass_blocks(p) * 256u, &sc)) != MRT_OK) return rc;
    MRT_HIP(launch_aov(s->dev, s->generic, p, d->first_sample, d->samples_pixel, dout, sc, s->err.as<uint32_t>(), st));
    if (mem != MRT_MEM_DEVICE) MRT_HIP(hipMemcpyAsync(aov, dout, bytes, hipMemcpyDeviceToHost, st));
    return check_err(s);
}

MRT_API int mrt_render_stop(mrt_scene* s) {
    if (!s) return fail(MRT_E_INVALID, "null scene");
    s->stop.store(true);
    return MRT_OK;
}

MRT_API int mrt_render_progress(const mrt_scene* s, uint32_t* samples_done) {
    if (!s || !samples_done) return fail(MRT_E_INVALID, "null argument");
    *samples_done = s->progress.load();
    return MRT_OK;
}

}  // extern "C"
