// Device-side trees of a GPU-built triangle BVH (declarations and layouts in
// mrt_walk_gpu.hpp).  Pass 1: interior flags -> exclusive scan (pair rank of
// each interior node, children of the reference's interior node i at
// ref[i].offset and + 1), parent links, each pair's depth by walking its
// parent links to the root (the root pair has depth 1), flags of the pairs
// at even depth -> exclusive scan (wide index).  Pass 2: one thread per
// interior node writes its child pair, its walk pair and, at even depth, its
// wide node.
#include "mrt_walk_gpu.hpp"

#include <algorithm>
#include <hipcub/hipcub.hpp>

#include "mrt_cwt.hpp"
#include "mrt_grid_reg.hpp"

namespace mrt {

namespace {

constexpr uint32_t kThreads = 256;

inline uint32_t blocks_for(size_t n) { return static_cast<uint32_t>((n + kThreads - 1) / kThreads); }

struct Carve {
    char* base;
    size_t off = 0;
    template <class T>
    T* take(size_t count) {
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += (count * sizeof(T) + 255) / 256 * 256;
        return p;
    }
};

// Scratch pieces (cnt reference nodes): interior flags and ranks (cnt + 1
// each: the last flag is 0, so the last rank is the pair count), parent
// links, per-pair depth, even-depth flags and wide indices (pairs + 1 each,
// sized for the worst case of cnt pairs), counters, root info, scan storage.
struct Pieces {
    uint32_t *flag, *rank, *parent, *depth, *even, *widx, *counters;
    TriTreeInfo* info;
    void* tmp;
    size_t tmp_bytes;
};

size_t carve(char* base, uint32_t cnt, size_t tmp_bytes, Pieces* p) {
    Carve c{base};
    p->flag = c.take<uint32_t>(cnt + 1);
    p->rank = c.take<uint32_t>(cnt + 1);
    p->parent = c.take<uint32_t>(cnt);
    p->depth = c.take<uint32_t>(cnt);
    p->even = c.take<uint32_t>(cnt + 1);
    p->widx = c.take<uint32_t>(cnt + 1);
    p->counters = c.take<uint32_t>(4);
    p->info = c.take<TriTreeInfo>(1);
    p->tmp = c.take<char>(tmp_bytes);
    p->tmp_bytes = tmp_bytes;
    return c.off;
}

__global__ __launch_bounds__(kThreads) void k_interior(const RefNode* __restrict__ ref, uint32_t cnt,
                                                       uint32_t* __restrict__ flag) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i <= cnt) flag[i] = i < cnt && ref[i].count == 0u ? 1u : 0u;
}

// Parent links of the interior nodes' children; counters[0] counts children
// outside the array (a malformed tree: the caller declines it).
__global__ __launch_bounds__(kThreads) void k_parent(const RefNode* __restrict__ ref, uint32_t cnt,
                                                     uint32_t* __restrict__ parent, uint32_t* counters) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= cnt || ref[i].count != 0u) return;
    const uint32_t c = ref[i].offset;
    if (c == 0u || c >= cnt - 1u) {
        atomicAdd(&counters[0], 1u);
        return;
    }
    parent[c] = i;
    parent[c + 1u] = i;
}

// Depth of each pair (number of interior nodes from the root down to it);
// counters[1] = the largest.  A walk longer than cnt steps (a cycle in a
// malformed tree) stops and is counted in counters[0].
__global__ __launch_bounds__(kThreads) void k_depth(const RefNode* __restrict__ ref, uint32_t cnt,
                                                    const uint32_t* __restrict__ rank,
                                                    const uint32_t* __restrict__ parent, uint32_t* __restrict__ depth,
                                                    uint32_t* __restrict__ even, uint32_t* counters) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= cnt || ref[i].count != 0u) return;
    uint32_t d = 1u, j = i;
    while (j != 0u && d <= cnt) {
        j = parent[j];
        ++d;
    }
    if (j != 0u) atomicAdd(&counters[0], 1u);
    const uint32_t g = rank[i];
    depth[g] = d;
    even[g] = (d & 1u) == 0u ? 1u : 0u;
    atomicMax(&counters[1], d);
}

__global__ void k_info(const RefNode* __restrict__ ref, const uint32_t* __restrict__ rank,
                       const uint32_t* __restrict__ widx, uint32_t cnt, const uint32_t* __restrict__ counters,
                       TriTreeInfo* info) {
    if (threadIdx.x != 0u) return;
    TriTreeInfo t{};
    t.pairs = rank[cnt];
    t.wide = widx[t.pairs];
    t.depth = counters[1];
    t.wide_levels = t.depth >= 2u ? (t.depth & ~1u) / 2u + 1u : 1u;
    t.root = ref[0];
    // (A leaf root, or children outside the array, leave the children zero;
    // the host declines both.)
    if (t.root.count == 0u && t.root.offset + 1u < cnt) {
        for (int k = 0; k < 2; ++k) {
            const uint32_t c = t.root.offset + static_cast<uint32_t>(k);
            t.child[k] = ref[c];
            t.child_wide[k] = t.child[k].count == 0u ? widx[rank[c]] : 0xFFFFFFFFu;
        }
    }
    *info = t;
}

__device__ __forceinline__ uint32_t leaf_word(uint32_t first, uint32_t count) {
    return (count << kSlotBits) | first;
}

struct Slot {
    Box box;
    uint32_t ref, word;
};

// A child of a wide node: a leaf as its leaf word, an interior node as its
// wide node.
__device__ __forceinline__ Slot wide_slot(const RefNode& n, uint32_t id, const uint32_t* rank, const uint32_t* widx,
                                          uint32_t wide_base) {
    if (n.count != 0u) return Slot{n.box, n.offset, leaf_word(n.offset, n.count)};
    return Slot{n.box, wide_base + widx[rank[id]], 0u};
}

__global__ __launch_bounds__(kThreads) void k_write(const RefNode* __restrict__ ref, uint32_t cnt,
                                                    const uint32_t* __restrict__ rank,
                                                    const uint32_t* __restrict__ depth,
                                                    const uint32_t* __restrict__ widx, uint32_t pair_base,
                                                    uint32_t wide_base, BNode* __restrict__ pairs,
                                                    BNode* __restrict__ walk, float4* __restrict__ wide) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= cnt || ref[i].count != 0u) return;
    const uint32_t g = rank[i];
    const uint32_t c = ref[i].offset;
    const RefNode L = ref[c], R = ref[c + 1u];
    BNode p;
    p.q0 = make_float4(L.box.mn.x, L.box.mn.y, L.box.mn.z, L.box.mx.x);
    p.q1 = make_float4(L.box.mx.y, L.box.mx.z, R.box.mn.x, R.box.mn.y);
    p.q2 = make_float4(R.box.mn.z, R.box.mx.x, R.box.mx.y, R.box.mx.z);
    const uint32_t lr = L.count ? L.offset : rank[c], rr = R.count ? R.offset : rank[c + 1u];
    if (pairs) {
        p.q3 = make_uint4(lr, L.count, rr, R.count);
        pairs[g] = p;
    }
    if (walk) {
        p.q3 = make_uint4(L.count ? lr : pair_base + lr, L.count ? leaf_word(lr, L.count) : 0u,
                          R.count ? rr : pair_base + rr, R.count ? leaf_word(rr, R.count) : 0u);
        walk[g] = p;
    }
    if (!wide || (depth[g] & 1u) != 0u) return;
    // Slots: each leaf child as it is, each interior child replaced by its
    // two children (left first).
    Slot s[4];
    int n = 0;
    const RefNode kids[2] = {L, R};
    for (int k = 0; k < 2; ++k) {
        const RefNode& ch = kids[k];
        if (ch.count != 0u) {
            s[n++] = Slot{ch.box, ch.offset, leaf_word(ch.offset, ch.count)};
        } else {
            const uint32_t gc = ch.offset;
            s[n++] = wide_slot(ref[gc], gc, rank, widx, wide_base);
            s[n++] = wide_slot(ref[gc + 1u], gc + 1u, rank, widx, wide_base);
        }
    }
    const float nan = __uint_as_float(0x7FC00000u);
    float v[6][4];
    uint32_t r[8];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool used = k < n;
        v[0][k] = used ? s[k].box.mn.x : nan;
        v[1][k] = used ? s[k].box.mn.y : nan;
        v[2][k] = used ? s[k].box.mn.z : nan;
        v[3][k] = used ? s[k].box.mx.x : nan;
        v[4][k] = used ? s[k].box.mx.y : nan;
        v[5][k] = used ? s[k].box.mx.z : nan;
        r[2 * k] = used ? s[k].ref : 0u;
        r[2 * k + 1] = used ? s[k].word : 0u;
    }
    float4* q = wide + 8ull * widx[g];
#pragma unroll
    for (int j = 0; j < 6; ++j) q[j] = make_float4(v[j][0], v[j][1], v[j][2], v[j][3]);
    q[6] = make_float4(__uint_as_float(r[0]), __uint_as_float(r[1]), __uint_as_float(r[2]), __uint_as_float(r[3]));
    q[7] = make_float4(__uint_as_float(r[4]), __uint_as_float(r[5]), __uint_as_float(r[6]), __uint_as_float(r[7]));
}

__global__ __launch_bounds__(kThreads) void k_pack_triangles(const float4* __restrict__ raw, uint32_t n,
                                                             float4* __restrict__ geo, float4* __restrict__ shade) {
    static_assert(sizeof(mrt_triangle) == 7 * sizeof(float4), "mrt_triangle is 112 bytes");
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    // mrt_triangle: ac[3] ab[3] a[3] na[3] nb[3] nc[3] ca[3] cb[3] cc[3] material_id.
    float f[28];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        const float4 q = raw[7ull * i + k];
        f[4 * k] = q.x, f[4 * k + 1] = q.y, f[4 * k + 2] = q.z, f[4 * k + 3] = q.w;
    }
    const float id = __uint_as_float(MRT_MAKE_ID(MRT_CAT_TRIANGLE, i));
    float4* g = geo + 3ull * i;
    g[0] = make_float4(f[6], f[7], f[8], f[3]);
    g[1] = make_float4(f[4], f[5], f[0], f[1]);
    g[2] = make_float4(f[2], id, 0.0f, 0.0f);
    float4* sh = shade + 5ull * i;
    sh[0] = make_float4(f[9], f[10], f[11], f[12]);
    sh[1] = make_float4(f[13], f[14], f[15], f[16]);
    sh[2] = make_float4(f[17], f[18], f[19], f[20]);
    sh[3] = make_float4(f[21], f[22], f[23], f[24]);
    sh[4] = make_float4(f[25], f[26], f[27], 0.0f);
}

__global__ __launch_bounds__(kThreads) void k_tri_boxes(const float4* __restrict__ geo, uint32_t n,
                                                        float* __restrict__ boxes6) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const float4 a = geo[3u * i], b = geo[3u * i + 1u], c = geo[3u * i + 2u];
    const Box bx = triangle_box_of(f3{a.x, a.y, a.z}, f3{a.w, b.x, b.y}, f3{b.z, b.w, c.x});
    float* o = boxes6 + 6ull * i;
    o[0] = bx.mn.x;
    o[1] = bx.mn.y;
    o[2] = bx.mn.z;
    o[3] = bx.mx.x;
    o[4] = bx.mx.y;
    o[5] = bx.mx.z;
}

// Compressed wide walk tree (layout and exactness in mrt_cwt.hpp): one
// thread per wide node writes its quantized node and its leaves' records.
// Slots with a finite box are quantized on a per-node, per-axis grid: origin
// the lowest bound of the node's finite slots, step a power of two with 255
// steps reaching the highest bound; each bound is rounded outward on the grid
// and then moved until the dequantized float contains the exact bound (a
// slot that cannot be is marked open).  Leaf words that do not fit the slot
// range write no record.
__device__ __forceinline__ uint32_t cw_exp_for(float lo, float hi) {
    const float d = (hi - lo) / 255.0f;
    uint32_t e = d > 0.0f ? ((__float_as_uint(d) >> 23) & 0xFFu) + 1u : 1u;
    e = e < 1u ? 1u : e;
    while (e < 254u && !(cw_deq(lo, 255u, cw_scale(e)) >= hi)) ++e;
    return e;
}

__global__ __launch_bounds__(kThreads) void k_cw_build(const float4* __restrict__ wide, uint32_t n_wide,
                                                       const float4* __restrict__ tri_geo, uint32_t n_slots,
                                                       uint4* __restrict__ nodes, float4* __restrict__ recs) {
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n_wide) return;
    const float4* q = wide + 8ull * i;
    const float4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4], q5 = q[5], q6 = q[6], q7 = q[7];
    const float mn[3][4] = {{q0.x, q0.y, q0.z, q0.w}, {q1.x, q1.y, q1.z, q1.w}, {q2.x, q2.y, q2.z, q2.w}};
    const float mx[3][4] = {{q3.x, q3.y, q3.z, q3.w}, {q4.x, q4.y, q4.z, q4.w}, {q5.x, q5.y, q5.z, q5.w}};
    const uint32_t ref[4] = {__float_as_uint(q6.x), __float_as_uint(q6.z), __float_as_uint(q7.x),
                             __float_as_uint(q7.z)};
    const uint32_t word[4] = {__float_as_uint(q6.y), __float_as_uint(q6.w), __float_as_uint(q7.y),
                              __float_as_uint(q7.w)};
    uint32_t kinds = 0u, open = 0u, finite = 0u, outRef[4];
    for (int k = 0; k < 4; ++k) {
        bool allNan = true, fin = true;
        for (int a = 0; a < 3; ++a) {
            allNan = allNan && isnan(mn[a][k]) && isnan(mx[a][k]);
            fin = fin && isfinite(mn[a][k]) && isfinite(mx[a][k]);
        }
        const uint32_t kind = word[k] != 0u ? kCwLeaf : (allNan ? kCwEmpty : kCwInterior);
        kinds |= kind << (2 * k);
        outRef[k] = kind == kCwLeaf ? word[k] : (kind == kCwInterior ? ref[k] : 0u);
        if (kind != kCwEmpty && fin) finite |= 1u << k;
        if (kind != kCwEmpty && !fin) open |= 1u << k;
    }
    float org[3] = {0.0f, 0.0f, 0.0f};
    uint32_t ex[3] = {127u, 127u, 127u}, qlo[3] = {0u, 0u, 0u}, qhi[3] = {0u, 0u, 0u};
    for (int a = 0; a < 3; ++a) {
        float lo = INFINITY, hi = -INFINITY;
        for (int k = 0; k < 4; ++k) {
            if (!(finite >> k & 1u)) continue;
            lo = fminf(lo, fminf(mn[a][k], mx[a][k]));
            hi = fmaxf(hi, fmaxf(mn[a][k], mx[a][k]));
        }
        if (!(lo <= hi)) continue;  // no finite slot
        const uint32_t e = cw_exp_for(lo, hi);
        const float sc = cw_scale(e);
        org[a] = lo;
        ex[a] = e;
        for (int k = 0; k < 4; ++k) {
            if (!(finite >> k & 1u)) continue;
            const float bl = fminf(mn[a][k], mx[a][k]), bh = fmaxf(mn[a][k], mx[a][k]);
            uint32_t l = static_cast<uint32_t>(fminf(fmaxf(floorf((bl - lo) / sc), 0.0f), 255.0f));
            while (l > 0u && cw_deq(lo, l, sc) > bl) --l;
            uint32_t h = static_cast<uint32_t>(fminf(fmaxf(ceilf((bh - lo) / sc), 0.0f), 255.0f));
            while (h < 255u && cw_deq(lo, h, sc) < bh) ++h;
            if (!(cw_deq(lo, l, sc) <= bl && cw_deq(lo, h, sc) >= bh)) {
                finite &= ~(1u << k);
                open |= 1u << k;
                continue;
            }
            qlo[a] |= l << (8 * k);
            qhi[a] |= h << (8 * k);
        }
    }
    // Open slots keep zero bytes (their test ignores them).
    for (int a = 0; a < 3; ++a)
        for (int k = 0; k < 4; ++k)
            if (open >> k & 1u) {
                qlo[a] &= ~(0xFFu << (8 * k));
                qhi[a] &= ~(0xFFu << (8 * k));
            }
    uint4* o = nodes + 4ull * i;
    o[0] = make_uint4(__float_as_uint(org[0]), __float_as_uint(org[1]), __float_as_uint(org[2]),
                      ex[0] | (ex[1] << 8) | (ex[2] << 16) | (kinds << 24));
    o[1] = make_uint4(qlo[0], qlo[1], qlo[2], qhi[0]);
    o[2] = make_uint4(qhi[1], qhi[2], outRef[0], outRef[1]);
    o[3] = make_uint4(outRef[2], outRef[3], open, 0u);
    // Leaf records: the exact box (as the reference's leaf has it) and two
    // triangles per record.
    for (int k = 0; k < 4; ++k) {
        if (word[k] == 0u) continue;
        const uint32_t slot = word[k] & kSlotMask, cnt = word[k] >> kSlotBits;
        if (cnt == 0u || slot >= n_slots || cnt > n_slots - slot) continue;
        for (uint32_t j = 0; 2u * j < cnt; ++j) {
            float4* r = recs + 8ull * (slot + j);
            r[0] = make_float4(mn[0][k], mn[1][k], mn[2][k], mx[0][k]);
            r[1] = make_float4(mx[1][k], mx[2][k], 0.0f, 0.0f);
            const float4* t0 = tri_geo + 3ull * (slot + 2u * j);
            r[2] = t0[0];
            r[3] = t0[1];
            r[4] = t0[2];
            if (2u * j + 1u < cnt) {
                r[5] = t0[3];
                r[6] = t0[4];
                r[7] = t0[5];
            } else {
                r[5] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                r[6] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                r[7] = make_float4(0.0f, __uint_as_float(MRT_ID_NONE), 0.0f, 0.0f);
            }
        }
    }
}

// A compressed node copied into a render server's arena: leaf words + tri_base
// (their slot field), interior references + wide_base.
__global__ __launch_bounds__(kThreads) void k_rebase_cw(const uint4* __restrict__ src, size_t n,
                                                        uint4* __restrict__ dst, uint32_t tri_base,
                                                        uint32_t wide_base) {
    const size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint4* q = src + 4 * i;
    uint4* o = dst + 4 * i;
    const uint4 n0 = q[0], n1 = q[1];
    uint4 n2 = q[2], n3 = q[3];
    const uint32_t kinds = n0.w >> 24;
    uint32_t r[4] = {n2.z, n2.w, n3.x, n3.y};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t kind = (kinds >> (2 * k)) & 3u;
        r[k] = kind == kCwLeaf ? r[k] + tri_base : (kind == kCwInterior ? r[k] + wide_base : 0u);
    }
    n2.z = r[0];
    n2.w = r[1];
    n3.x = r[2];
    n3.y = r[3];
    o[0] = n0;
    o[1] = n1;
    o[2] = n2;
    o[3] = n3;
}

__global__ __launch_bounds__(kThreads) void k_rebase_walk(const BNode* __restrict__ src, size_t n,
                                                          BNode* __restrict__ dst, uint32_t tri_base,
                                                          uint32_t walk_base) {
    const size_t i = static_cast<size_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n) return;
    BNode p = src[i];
    const bool nanL = isnan(p.q0.x), nanR = isnan(p.q1.z);
    if (p.q3.y != 0u) {
        p.q3.x += tri_base;
        p.q3.y += tri_base;
    } else if (!nanL) {
        p.q3.x += walk_base;
    }
    if (p.q3.w != 0u) {
        p.q3.z += tri_base;
        p.q3.w += tri_base;
    } else if (!nanR) {
        p.q3.z += walk_base;
    }
    dst[i] = p;
}

#define WALK_HIP(expr)                   \
    do {                                 \
        const hipError_t e_ = (expr);    \
        if (e_ != hipSuccess) return e_; \
    } while (0)

size_t scan_bytes(uint32_t cnt, hipStream_t st) {
    size_t b = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, b, static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                         static_cast<int>(cnt + 1), st) != hipSuccess)
        return 0;
    return b;
}

}  // namespace

WalkGpuScratch::~WalkGpuScratch() {
    if (p) (void)hipFree(p);
}

hipError_t gpu_tri_tree_plan(const RefNode* ref, uint32_t cnt, WalkGpuScratch* sc, TriTreeInfo* info,
                             hipStream_t st) {
    if (cnt < 3u || cnt >= (1u << 30)) return hipErrorInvalidValue;
    const size_t tmp = scan_bytes(cnt, st);
    if (tmp == 0) return hipErrorInvalidValue;
    Pieces P;
    const size_t need = carve(nullptr, cnt, tmp, &P);
    if (sc->bytes < need) {
        if (sc->p) WALK_HIP(hipFree(sc->p));
        sc->p = nullptr;
        sc->bytes = 0;
        WALK_HIP(hipMalloc(&sc->p, need));
        sc->bytes = need;
    }
    carve(static_cast<char*>(sc->p), cnt, tmp, &P);
    WALK_HIP(hipMemsetAsync(P.parent, 0, static_cast<size_t>(cnt) * sizeof(uint32_t), st));
    WALK_HIP(hipMemsetAsync(P.counters, 0, 4 * sizeof(uint32_t), st));
    k_interior<<<blocks_for(cnt + 1), kThreads, 0, st>>>(ref, cnt, P.flag);
    size_t t = tmp;
    WALK_HIP(hipcub::DeviceScan::ExclusiveSum(P.tmp, t, P.flag, P.rank, static_cast<int>(cnt + 1), st));
    k_parent<<<blocks_for(cnt), kThreads, 0, st>>>(ref, cnt, P.parent, P.counters);
    // Every pair's even flag is written by k_depth; the one past the last
    // pair must read 0 for the scan's total.  Pairs never exceed cnt, so the
    // whole array is cleared first.
    WALK_HIP(hipMemsetAsync(P.even, 0, (static_cast<size_t>(cnt) + 1) * sizeof(uint32_t), st));
    k_depth<<<blocks_for(cnt), kThreads, 0, st>>>(ref, cnt, P.rank, P.parent, P.depth, P.even, P.counters);
    t = tmp;
    WALK_HIP(hipcub::DeviceScan::ExclusiveSum(P.tmp, t, P.even, P.widx, static_cast<int>(cnt + 1), st));
    k_info<<<1, 64, 0, st>>>(ref, P.rank, P.widx, cnt, P.counters, P.info);
    WALK_HIP(hipGetLastError());
    uint32_t bad = 0;
    WALK_HIP(hipMemcpyAsync(info, P.info, sizeof(TriTreeInfo), hipMemcpyDeviceToHost, st));
    WALK_HIP(hipMemcpyAsync(&bad, P.counters, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    WALK_HIP(hipStreamSynchronize(st));
    if (bad != 0u || info->root.count != 0u || info->pairs == 0u) return hipErrorInvalidValue;
    return hipSuccess;
}

hipError_t gpu_tri_tree_write(const RefNode* ref, uint32_t cnt, const WalkGpuScratch& sc, const TriTreeInfo& info,
                              uint32_t pair_base, uint32_t wide_base, BNode* pairs, BNode* walk, float4* wide,
                              hipStream_t st) {
    if (info.pairs == 0u || info.pairs > cnt) return hipErrorInvalidValue;
    const size_t tmp = scan_bytes(cnt, st);
    Pieces P;
    if (carve(nullptr, cnt, tmp, &P) > sc.bytes) return hipErrorInvalidValue;
    carve(static_cast<char*>(sc.p), cnt, tmp, &P);
    k_write<<<blocks_for(cnt), kThreads, 0, st>>>(ref, cnt, P.rank, P.depth, P.widx, pair_base, wide_base, pairs, walk,
                                                  wide);
    return hipGetLastError();
}

hipError_t gpu_rebase_cw(const uint4* src, size_t n, uint4* dst, uint32_t tri_base, uint32_t wide_base,
                         hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_rebase_cw<<<blocks_for(n), kThreads, 0, st>>>(src, n, dst, tri_base, wide_base);
    return hipGetLastError();
}

hipError_t gpu_rebase_walk(const BNode* src, size_t n, BNode* dst, uint32_t tri_base, uint32_t walk_base,
                           hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_rebase_walk<<<blocks_for(n), kThreads, 0, st>>>(src, n, dst, tri_base, walk_base);
    return hipGetLastError();
}

hipError_t gpu_pack_triangles(const mrt_triangle* tris, uint32_t n, float4* geo, float4* shade, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_pack_triangles<<<blocks_for(n), kThreads, 0, st>>>(reinterpret_cast<const float4*>(tris), n, geo, shade);
    return hipGetLastError();
}

hipError_t gpu_compress_walk(const float4* wide, uint32_t n_wide, const float4* tri_geo, uint32_t n_slots,
                             uint4* nodes, float4* recs, hipStream_t st) {
    if (n_wide == 0u) return hipSuccess;
    k_cw_build<<<blocks_for(n_wide), kThreads, 0, st>>>(wide, n_wide, tri_geo, n_slots, nodes, recs);
    return hipGetLastError();
}

hipError_t gpu_tri_boxes(const float4* tri_geo, uint32_t n, float* boxes6, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_tri_boxes<<<blocks_for(n), kThreads, 0, st>>>(tri_geo, n, boxes6);
    return hipGetLastError();
}

}  // namespace mrt
