// Host-side builders: the reference BVH builder, the conversion to 64-byte
// child pairs, the light BVH, primitive boxes and box-overlap tests.
#include "mrt_build.hpp"
#include "mrt_grid_reg.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>

namespace mrt {

namespace {

// getSplitIndex_SAH, BVH.hpp:274-311.  right[i] holds the area of
// boxes[i + L .. N-1] (the reference's indexing, one past the split point).
uint32_t split_sah(const Box* bx, uint32_t N, uint32_t L, std::vector<float>& left, std::vector<float>& right) {
    left.assign(N - L, 0.0f);
    right.assign(N - L, 0.0f);
    Box lb = bx[0];
    left[0] = surface_area(lb);
    for (uint32_t i = 1; i < N - L; ++i) {
        lb = surrounding(lb, bx[i]);
        left[i] = surface_area(lb);
    }
    Box rb = bx[N - 1];
    right[N - L - 1] = surface_area(rb);
    for (uint32_t i = N - 2; i > L; --i) {
        rb = surrounding(rb, bx[i]);
        right[i - L] = surface_area(rb);
    }
    uint32_t split = L;
    float best = static_cast<float>(L) * left[L - 1] + static_cast<float>(N - L) * right[L - 1];
    for (uint32_t i = L; i < N - L; ++i) {
        const float sl = static_cast<float>(i + 1) * left[i];
        const float sr = static_cast<float>(N - (i + 1)) * right[i];
        const float sah = sl + sr;
        if (sah < best) {
            split = i + 1;
            best = sah;
        }
    }
    return split;
}

void to_pairs_impl(BvhOut* out) {
    const std::vector<RefNode>& ref = out->ref;
    out->pairs.clear();
    out->depth = 0;
    out->root = ref[0].box;
    if (ref[0].count > 0) {
        out->root_ref = ref[0].offset;
        out->root_count = ref[0].count;
        return;
    }
    out->root_ref = 0;
    out->root_count = 0;
    // Interior reference node i becomes pair g[i], numbered in reference order.
    std::vector<uint32_t> g(ref.size(), 0xFFFFFFFFu);
    uint32_t k = 0;
    for (size_t i = 0; i < ref.size(); ++i)
        if (ref[i].count == 0) g[i] = k++;
    out->pairs.resize(k);
    for (size_t i = 0; i < ref.size(); ++i) {
        if (ref[i].count != 0) continue;
        const RefNode& L = ref[ref[i].offset];
        const RefNode& R = ref[ref[i].offset + 1];
        BNode& p = out->pairs[g[i]];
        p.q0 = make_float4(L.box.mn.x, L.box.mn.y, L.box.mn.z, L.box.mx.x);
        p.q1 = make_float4(L.box.mx.y, L.box.mx.z, R.box.mn.x, R.box.mn.y);
        p.q2 = make_float4(R.box.mn.z, R.box.mx.x, R.box.mx.y, R.box.mx.z);
        const uint32_t lref = L.count > 0 ? L.offset : g[ref[i].offset];
        const uint32_t rref = R.count > 0 ? R.offset : g[ref[i].offset + 1];
        p.q3 = make_uint4(lref, L.count, rref, R.count);
    }
    // Longest chain of interior pairs (bounds the traversal stack).
    std::vector<std::pair<uint32_t, uint32_t>> st{{0u, 1u}};
    while (!st.empty()) {
        const auto [node, d] = st.back();
        st.pop_back();
        out->depth = std::max(out->depth, d);
        const uint4 q = out->pairs[node].q3;
        if (q.y == 0) st.push_back({q.x, d + 1});
        if (q.w == 0) st.push_back({q.z, d + 1});
    }
}

}  // namespace

void build_bvh_reference(const std::vector<Box>& boxes, uint32_t leaf, BvhOut* out) {
    out->ref.clear();
    const uint32_t n = static_cast<uint32_t>(boxes.size());
    if (n == 0) {
        out->ref.push_back(RefNode{Box{f3{0, 0, 0}, f3{0, 0, 0}}, 0, 0});
        out->pairs.clear();
        out->root = out->ref[0].box;
        out->root_ref = 0;
        out->root_count = 0;
        out->depth = 0;
        return;
    }
    std::vector<RefNode>& nodes = out->ref;
    nodes.assign(static_cast<size_t>(n) * 2 - 1, RefNode{Box{}, 0, 0});
    std::vector<float> la, ra;
    uint32_t id = 0, begin = 0, end = n, maxId = 0;
    // Explicit stacks as in BVH.hpp:93-104; entry 0 is the sentinel.
    std::vector<uint32_t> sId{0}, sBegin{0}, sEnd{0};
    do {
        RefNode& node = nodes[id];
        node.box = boxes[begin];
        for (uint32_t i = begin + 1; i < end; ++i) node.box = surrounding(boxes[i], node.box);
        const uint32_t count = end - begin;
        if (count <= leaf) {
            node.offset = begin;
            node.count = count;
            id = sId.back();
            sId.pop_back();
            begin = sBegin.back();
            sBegin.pop_back();
            end = sEnd.back();
            sEnd.pop_back();
        } else {
            const uint32_t left = maxId + 1, right = left + 1;
            node.offset = left;
            node.count = 0;
            maxId = std::max(right, maxId);
            const uint32_t split = count <= 2 * leaf ? leaf : split_sah(boxes.data() + begin, count, leaf, la, ra);
            sId.push_back(right);
            sBegin.push_back(begin + split);
            sEnd.push_back(end);
            id = left;
            end = begin + split;
        }
    } while (!sId.empty());
    nodes.resize(static_cast<size_t>(maxId) + 1);
    to_pairs_impl(out);
}

void bvh_pairs_from_ref(BvhOut* out) { to_pairs_impl(out); }

Box triangle_box(const mrt_triangle& t) { return triangle_box_of(ld3(t.a), ld3(t.ab), ld3(t.ac)); }

Box sphere_box(const mrt_sphere& s) { return sphere_box_of(ld3(s.center), s.sq_radius); }

Box plane_box(const mrt_plane& p) { return plane_box_of(ld3(p.normal), ld3(p.point)); }

bool triangle_intersect_box(const mrt_triangle& t, const Box& box) {
    return triangle_box_overlap(ld3(t.a), ld3(t.ab), ld3(t.ac), box);
}

bool sphere_intersect_box(const mrt_sphere& s, const Box& box) {
    return sphere_box_overlap(ld3(s.center), s.sq_radius, box);
}

bool plane_intersect_box(const mrt_plane& p, const Box& box) {
    return plane_box_overlap(ld3(p.normal), ld3(p.point), box);
}

int32_t grid_size_for(uint32_t n) {
    auto bits = [](uint32_t v) {
        uint32_t c = 0;
        while (v > 0) { ++c; v >>= 1; }
        return c;
    };
    auto pow2 = [](uint32_t v) {
        --v;
        v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16;
        return v + 1;
    };
    const int32_t aux = n > 6 ? static_cast<int32_t>(pow2(bits(n))) : 2;
    const int32_t aux2 = static_cast<int32_t>(pow2(bits(static_cast<uint32_t>(aux))) / 2);
    return std::max(aux * aux2, aux);
}

void build_light_bvh(const std::vector<Box>& boxes, BvhOut* out, std::vector<uint32_t>* perm) {
    const uint32_t n = static_cast<uint32_t>(boxes.size());
    // Morton order of box centres, then the in-order builder with leaf size 2.
    Box all = boxes.empty() ? Box{} : boxes[0];
    for (const Box& b : boxes) all = surrounding(b, all);
    const f3 ext = sub(all.mx, all.mn);
    auto spread = [](uint32_t v) {
        v &= 0x3FFu;
        v = (v | (v << 16)) & 0x030000FFu;
        v = (v | (v << 8)) & 0x0300F00Fu;
        v = (v | (v << 4)) & 0x030C30C3u;
        v = (v | (v << 2)) & 0x09249249u;
        return v;
    };
    std::vector<uint32_t> code(n);
    for (uint32_t i = 0; i < n; ++i) {
        const f3 c = muls(add(boxes[i].mn, boxes[i].mx), 0.5f);
        auto q = [](float v, float lo, float e) {
            const float t = e > 0.0f ? (v - lo) / e : 0.0f;
            return static_cast<uint32_t>(std::min(1023.0f, std::max(0.0f, t * 1023.0f)));
        };
        code[i] = (spread(q(c.x, all.mn.x, ext.x)) << 2) | (spread(q(c.y, all.mn.y, ext.y)) << 1) |
                  spread(q(c.z, all.mn.z, ext.z));
    }
    perm->resize(n);
    std::iota(perm->begin(), perm->end(), 0u);
    std::stable_sort(perm->begin(), perm->end(), [&](uint32_t a, uint32_t b) { return code[a] < code[b]; });
    std::vector<Box> sorted(n);
    for (uint32_t i = 0; i < n; ++i) {
        Box b = boxes[(*perm)[i]];
        // Widen by a relative margin so rounding in the slab test can never
        // reject a point of the triangle.
        const f3 e = sub(b.mx, b.mn);
        const float m = 1e-4f * std::max(e.x, std::max(e.y, e.z)) + 1e-6f;
        const f3 ma = f3{m + 1e-6f * std::fabs(b.mn.x), m + 1e-6f * std::fabs(b.mn.y), m + 1e-6f * std::fabs(b.mn.z)};
        const f3 mb = f3{m + 1e-6f * std::fabs(b.mx.x), m + 1e-6f * std::fabs(b.mx.y), m + 1e-6f * std::fabs(b.mx.z)};
        b.mn = sub(b.mn, ma);
        b.mx = add(b.mx, mb);
        sorted[i] = b;
    }
    build_bvh_reference(sorted, 2u, out);
}

}  // namespace mrt

// ------------------------------------------------------------- wide SAH --
namespace mrt {

void put_wide_node(std::vector<float4>* out, uint32_t index, const WideSlot* sl, int n) {
    const float nan = std::numeric_limits<float>::quiet_NaN();
    if (out->size() < 8ull * (index + 1u)) out->resize(8ull * (index + 1u));
    float4* q = &(*out)[8ull * index];
    float v[6][4];
    uint32_t r[8];
    for (int k = 0; k < 4; ++k) {
        const bool used = k < n;
        const Box b = used ? sl[k].box : Box{f3{nan, nan, nan}, f3{nan, nan, nan}};
        v[0][k] = b.mn.x;
        v[1][k] = b.mn.y;
        v[2][k] = b.mn.z;
        v[3][k] = b.mx.x;
        v[4][k] = b.mx.y;
        v[5][k] = b.mx.z;
        r[2 * k] = used ? sl[k].ref : 0u;
        r[2 * k + 1] = used ? sl[k].word : 0u;
    }
    for (int j = 0; j < 6; ++j) q[j] = make_float4(v[j][0], v[j][1], v[j][2], v[j][3]);
    std::memcpy(&q[6], r, 16);
    std::memcpy(&q[7], r + 4, 16);
}

namespace {

struct BinNode {
    Box box;
    int32_t l = -1, r = -1;  // children (interior), -1 for a unit
    uint32_t unit = 0;
};

float axis_of(const f3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// Binary tree over the units: binned SAH (16 bins, every axis) on the unit
// centres, an index median where no bin split separates them.
void build_binary(const std::vector<WideSlot>& units, std::vector<BinNode>* nodes) {
    constexpr int kBins = 16;
    const uint32_t n = static_cast<uint32_t>(units.size());
    std::vector<uint32_t> idx(n);
    std::vector<f3> cen(n);
    for (uint32_t i = 0; i < n; ++i) {
        idx[i] = i;
        cen[i] = muls(add(units[i].box.mn, units[i].box.mx), 0.5f);
    }
    nodes->clear();
    nodes->reserve(2ull * n);
    nodes->push_back(BinNode{});
    struct Task {
        int32_t node;
        uint32_t b, e;
    };
    std::vector<Task> stack{{0, 0, n}};
    while (!stack.empty()) {
        const Task t = stack.back();
        stack.pop_back();
        Box box = units[idx[t.b]].box;
        Box cb{cen[idx[t.b]], cen[idx[t.b]]};
        for (uint32_t k = t.b + 1; k < t.e; ++k) {
            box = surrounding(box, units[idx[k]].box);
            cb = Box{vmin(cb.mn, cen[idx[k]]), vmax(cb.mx, cen[idx[k]])};
        }
        (*nodes)[t.node].box = box;
        if (t.e - t.b == 1) {
            (*nodes)[t.node].unit = idx[t.b];
            continue;
        }
        int bestAxis = -1, bestSplit = 0;
        float bestCost = std::numeric_limits<float>::infinity();
        for (int a = 0; a < 3; ++a) {
            const float lo = axis_of(cb.mn, a), ext = axis_of(cb.mx, a) - lo;
            if (!(ext > 0.0f)) continue;
            Box bb[kBins];
            uint32_t cnt[kBins] = {};
            for (uint32_t k = t.b; k < t.e; ++k) {
                const int bi = std::min(kBins - 1, static_cast<int>((axis_of(cen[idx[k]], a) - lo) * (kBins / ext)));
                bb[bi] = cnt[bi] ? surrounding(bb[bi], units[idx[k]].box) : units[idx[k]].box;
                ++cnt[bi];
            }
            float right[kBins];
            Box rb{};
            uint32_t rc = 0;
            for (int i = kBins - 1; i > 0; --i) {
                if (cnt[i]) {
                    rb = rc ? surrounding(rb, bb[i]) : bb[i];
                    rc += cnt[i];
                }
                right[i] = rc ? surface_area(rb) * static_cast<float>(rc) : 0.0f;
            }
            Box lb{};
            uint32_t lc = 0;
            for (int i = 1; i < kBins; ++i) {
                if (cnt[i - 1]) {
                    lb = lc ? surrounding(lb, bb[i - 1]) : bb[i - 1];
                    lc += cnt[i - 1];
                }
                if (lc == 0 || lc == t.e - t.b) continue;
                const float cost = surface_area(lb) * static_cast<float>(lc) + right[i];
                if (cost < bestCost) {
                    bestCost = cost;
                    bestAxis = a;
                    bestSplit = i;
                }
            }
        }
        uint32_t mid;
        if (bestAxis >= 0) {
            const float lo = axis_of(cb.mn, bestAxis), ext = axis_of(cb.mx, bestAxis) - lo;
            const auto it = std::partition(idx.begin() + t.b, idx.begin() + t.e, [&](uint32_t u) {
                return std::min(kBins - 1, static_cast<int>((axis_of(cen[u], bestAxis) - lo) * (kBins / ext))) <
                       bestSplit;
            });
            mid = static_cast<uint32_t>(it - idx.begin());
        } else {
            mid = (t.b + t.e) / 2u;  // coincident centres: any halving
        }
        if (mid == t.b || mid == t.e) mid = (t.b + t.e) / 2u;
        const int32_t l = static_cast<int32_t>(nodes->size());
        nodes->push_back(BinNode{});
        nodes->push_back(BinNode{});
        (*nodes)[t.node].l = l;
        (*nodes)[t.node].r = l + 1;
        stack.push_back(Task{l, t.b, mid});
        stack.push_back(Task{l + 1, mid, t.e});
    }
}

// Up to `limit` binary nodes covering node b's subtree: b's children, then the
// interior one with the largest surface area replaced by its children, and so
// on.
int expand(const std::vector<BinNode>& bn, int32_t b, int limit, int32_t* out) {
    if (bn[b].l < 0) {
        out[0] = b;
        return 1;
    }
    int n = 0;
    out[n++] = bn[b].l;
    out[n++] = bn[b].r;
    while (n < limit) {
        int best = -1;
        float bestArea = -1.0f;
        for (int k = 0; k < n; ++k) {
            if (bn[out[k]].l < 0) continue;
            const float a = surface_area(bn[out[k]].box);
            if (a > bestArea) {
                bestArea = a;
                best = k;
            }
        }
        if (best < 0) break;
        const int32_t c = out[best];
        out[best] = bn[c].l;
        out[n++] = bn[c].r;
    }
    return n;
}

}  // namespace

void build_wide_sah(const std::vector<WideSlot>& units, uint32_t base, int max_top, std::vector<float4>* out,
                    std::vector<WideSlot>* top, uint32_t* depth) {
    top->clear();
    *depth = 0;
    if (units.empty()) return;
    std::vector<BinNode> bn;
    build_binary(units, &bn);
    constexpr uint32_t kNone = 0xFFFFFFFFu;
    uint32_t next = base;
    // Slot of binary node c: a unit as its own leaf slot, an interior node as
    // a wide node to be written.
    struct Pend {
        int32_t bin;
        uint32_t index, level;
    };
    std::vector<Pend> todo;
    auto slot_of = [&](int32_t c, uint32_t level) {
        if (bn[c].l < 0) return units[bn[c].unit];
        const uint32_t index = next++;
        todo.push_back(Pend{c, index, level});
        return WideSlot{bn[c].box, index, 0u};
    };
    int32_t sel[4];
    const int nt = expand(bn, 0, std::max(1, std::min(4, max_top)), sel);
    for (int k = 0; k < nt; ++k) top->push_back(slot_of(sel[k], 1u));
    (void)kNone;
    while (!todo.empty()) {
        const Pend p = todo.back();
        todo.pop_back();
        *depth = std::max(*depth, p.level);
        const int n = expand(bn, p.bin, 4, sel);
        WideSlot sl[4];
        for (int k = 0; k < n; ++k) sl[k] = slot_of(sel[k], p.level + 1u);
        put_wide_node(out, p.index - base, sl, n);
    }
}

}  // namespace mrt
