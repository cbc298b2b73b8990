// GPU build of the reference BVH, level by level.
//
// The reference builder (BVH.hpp:64-155) never reorders primitives: a node
// covers a contiguous range [begin, end) of them in scene order; a range of
// at most `leaf` primitives is a leaf, one of at most 2 * leaf splits at
// begin + leaf, and a larger one at begin + s, s = 1 + the first index i in
// [leaf - 1, N - leaf - 1] minimising
//   (i + 1) * area(boxes[0..i]) + (N - i - 1) * R(i),
//   R(i) = i == 0 ? 0 : area(boxes[i + leaf .. N - 1])
// (getSplitIndex_SAH, BVH.hpp:274-311, with its indexing; N = end - begin,
// indices relative to begin).  The children of the interior node of
// pre-order rank k get ids 2k + 1 and 2k + 2 (the explicit-stack walk of
// BVH.hpp:93-104).
//
// Nodes of more than kSmallNode primitives are split level by level:
// segmented scans give every position's prefix and suffix box within its
// node, one thread per position evaluates the cost, and a 64-bit atomic
// minimum of (cost bits, index) per node keeps the first minimum (costs are
// non-negative, so their bit patterns order like their values; -0 becomes
// +0).  Subtrees of at most kSmallNode primitives are built by one thread
// each with the sequential algorithm.  A pre-order numbering pass then gives
// every node its id.
//
// Boxes are combined with the reference's rule for node boxes
// (node.box = surrounding(boxes[i], node.box)): on equal coordinates the
// later box's value stays, which fixes the sign of a zero.  For non-NaN
// values that rule is associative, so scans in any grouping reproduce the
// sequential boxes bit for bit; inputs with a NaN coordinate are left to
// the host.
#include "mrt_bvh_gpu.hpp"

#include <utility>
#include <vector>

#include "mrt_math.hpp"

namespace mrt {

namespace {

constexpr uint32_t kSmallNode = 64;
constexpr uint32_t kScanThreads = 256, kScanItems = 4, kScanBlock = kScanThreads * kScanItems;
constexpr uint32_t kMaxLevels = 1024;
constexpr uint32_t kLeafKind = 0, kSmallKind = 1, kBigKind = 2;

struct GBox {
    float v[6];
};

__device__ __forceinline__ GBox load_box(const float* b, uint32_t k) {
    const float2* p = reinterpret_cast<const float2*>(b + 6 * static_cast<size_t>(k));
    const float2 a = p[0], c = p[1], d = p[2];
    return GBox{{a.x, a.y, c.x, c.y, d.x, d.y}};
}

__device__ __forceinline__ Box to_box(const GBox& g) {
    return Box{f3{g.v[0], g.v[1], g.v[2]}, f3{g.v[3], g.v[4], g.v[5]}};
}

__device__ __forceinline__ GBox from_box(const Box& b) {
    return GBox{{b.mn.x, b.mn.y, b.mn.z, b.mx.x, b.mx.y, b.mx.z}};
}

// Union of an earlier run e and a later run l: surrounding(l, e).
__device__ __forceinline__ GBox join(const GBox& e, const GBox& l) {
    GBox r;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        r.v[a] = (e.v[a] < l.v[a]) ? e.v[a] : l.v[a];
        r.v[3 + a] = (l.v[3 + a] < e.v[3 + a]) ? e.v[3 + a] : l.v[3 + a];
    }
    return r;
}

// Combination in scan order (first, second); a reverse scan visits later
// indices first.
template <bool REV>
__device__ __forceinline__ GBox op(const GBox& first, const GBox& second) {
    return REV ? join(second, first) : join(first, second);
}

__device__ __forceinline__ GBox identity_box() {
    const float inf = __builtin_inff();
    return GBox{{inf, inf, inf, -inf, -inf, -inf}};
}

// Node pool of the level-by-level phase (structure of arrays).
struct Pool {
    uint32_t* begin;
    uint32_t* end;
    GBox* box;
    uint32_t* child;  // left child (right = left + 1) of a split node
    uint32_t* kind;
    unsigned long long* key;
    uint32_t* mid;
    uint32_t* ic;    // interior nodes in the subtree
    uint32_t* rank;  // pre-order rank among interior nodes
    uint32_t* id;    // reference node id
};

struct ScanBuf {
    const float* boxes;
    const int32_t* seg;
    GBox* out;
    uint32_t n;
    GBox* blk_box;
    uint32_t* blk_head;
    uint32_t* blk_first;
    GBox* blk_carry;
};

// Segmented inclusive scan, phase 1: within blocks of kScanBlock positions.
// A position starts a segment when its node differs from the previous
// position's; positions of no node (seg < 0) are segments of their own.
template <bool REV>
__global__ __launch_bounds__(kScanThreads) void k_scan_local(ScanBuf s) {
    __shared__ GBox sb[kScanThreads];
    __shared__ uint32_t sf[kScanThreads];
    __shared__ uint32_t first_head;
    const uint32_t t = threadIdx.x, b = blockIdx.x;
    if (t == 0) first_head = kScanBlock;
    __syncthreads();
    const uint32_t p0 = b * kScanBlock + t * kScanItems;
    GBox v[kScanItems];
    bool head[kScanItems];
    int32_t sg[kScanItems];
    bool anyHead = false;
#pragma unroll
    for (uint32_t i = 0; i < kScanItems; ++i) {
        const uint32_t p = p0 + i;
        if (p >= s.n) {
            head[i] = true;
            sg[i] = -1;
            v[i] = identity_box();
        } else {
            const uint32_t k = REV ? s.n - 1 - p : p;
            sg[i] = s.seg[k];
            const int32_t prev = p > 0 ? s.seg[REV ? k + 1 : k - 1] : -2;
            head[i] = p == 0 || sg[i] < 0 || sg[i] != prev;
            const GBox x = sg[i] >= 0 ? load_box(s.boxes, k) : identity_box();
            v[i] = (head[i] || i == 0) ? x : op<REV>(v[i - 1], x);
        }
        if (head[i]) {
            anyHead = true;
            atomicMin(&first_head, t * kScanItems + i);
        }
    }
    sb[t] = v[kScanItems - 1];
    sf[t] = anyHead ? 1u : 0u;
    __syncthreads();
    for (uint32_t d = 1; d < kScanThreads; d <<= 1) {
        GBox pb = identity_box(), cb = sb[t];
        uint32_t pf = 0, cf = sf[t];
        if (t >= d) {
            pb = sb[t - d];
            pf = sf[t - d];
        }
        __syncthreads();
        if (t >= d) {
            if (!cf) cb = op<REV>(pb, cb);
            sb[t] = cb;
            sf[t] = cf | pf;
        }
        __syncthreads();
    }
    // The run that reaches this thread's first positions from earlier
    // threads of the block.
    if (t > 0 && !head[0]) {
        const GBox c = sb[t - 1];
#pragma unroll
        for (uint32_t i = 0; i < kScanItems; ++i) {
            if (head[i]) break;
            v[i] = op<REV>(c, v[i]);
        }
    }
#pragma unroll
    for (uint32_t i = 0; i < kScanItems; ++i) {
        const uint32_t p = p0 + i;
        if (p < s.n && sg[i] >= 0) s.out[REV ? s.n - 1 - p : p] = v[i];
    }
    if (t == kScanThreads - 1) {
        s.blk_box[b] = sb[t];
        s.blk_head[b] = sf[t];
        s.blk_first[b] = first_head;
    }
}

// Phase 2: the run entering each block from the blocks before it (one
// workgroup; thread t walks a contiguous chunk of blocks).
template <bool REV>
__global__ __launch_bounds__(256) void k_scan_blocks(ScanBuf s, uint32_t nb) {
    __shared__ GBox sb[256];
    __shared__ uint32_t sf[256];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nb + 255u) / 256u;
    const uint32_t c0 = t * per < nb ? t * per : nb, c1 = c0 + per < nb ? c0 + per : nb;
    GBox acc = identity_box();
    uint32_t f = 0;
    for (uint32_t b = c0; b < c1; ++b) {
        acc = s.blk_head[b] ? s.blk_box[b] : op<REV>(acc, s.blk_box[b]);
        f |= s.blk_head[b];
    }
    sb[t] = acc;
    sf[t] = f;
    __syncthreads();
    for (uint32_t d = 1; d < 256u; d <<= 1) {
        GBox pb = identity_box(), cb = sb[t];
        uint32_t pf = 0, cf = sf[t];
        if (t >= d) {
            pb = sb[t - d];
            pf = sf[t - d];
        }
        __syncthreads();
        if (t >= d) {
            if (!cf) cb = op<REV>(pb, cb);
            sb[t] = cb;
            sf[t] = cf | pf;
        }
        __syncthreads();
    }
    GBox run = t > 0 ? sb[t - 1] : identity_box();
    for (uint32_t b = c0; b < c1; ++b) {
        s.blk_carry[b] = run;
        run = s.blk_head[b] ? s.blk_box[b] : op<REV>(run, s.blk_box[b]);
    }
}

// Phase 3: positions of a block before its first segment start continue the
// run of the blocks before it.
template <bool REV>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(ScanBuf s) {
    const uint32_t b = blockIdx.x;
    if (b == 0) return;
    const uint32_t fh = s.blk_first[b];
    const GBox c = s.blk_carry[b];
#pragma unroll
    for (uint32_t i = 0; i < kScanItems; ++i) {
        const uint32_t off = threadIdx.x * kScanItems + i;
        const uint32_t p = b * kScanBlock + off;
        if (off >= fh || p >= s.n) break;
        const uint32_t k = REV ? s.n - 1 - p : p;
        s.out[k] = op<REV>(c, s.out[k]);
    }
}

// Split costs of every position of the nodes being split; the first
// minimum per node survives in key.
__global__ __launch_bounds__(256) void k_sah(const int32_t* __restrict__ seg, const GBox* __restrict__ pre,
                                             const GBox* __restrict__ suf, Pool P, uint32_t n, uint32_t L) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    int32_t s = -1;
    unsigned long long key = ~0ull;
    if (k < n) {
        s = seg[k];
        if (s >= 0) {
            const uint32_t b = P.begin[s], N = P.end[s] - b, i = k - b;
            if (i + 1 >= L && i + L < N) {
                const float left = surface_area(to_box(pre[k]));
                const float right = i == 0 ? 0.0f : surface_area(to_box(suf[k + L]));
                const float sl = static_cast<float>(i + 1) * left;
                const float sr = static_cast<float>(N - (i + 1)) * right;
                const float cost = (sl + sr) + 0.0f;
                uint32_t bits = __float_as_uint(cost);
                // A NaN first candidate: no later cost compares below it.
                if (i + 1 == L && cost != cost) bits = 0u;
                key = (static_cast<unsigned long long>(bits) << 32) | i;
            }
        }
    }
    const int32_t s0 = __shfl(s, 0);
    if (__all(s == s0)) {
        if (s0 < 0) return;
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t lo = __shfl_xor(static_cast<uint32_t>(key), off);
            const uint32_t hi = __shfl_xor(static_cast<uint32_t>(key >> 32), off);
            const unsigned long long o = (static_cast<unsigned long long>(hi) << 32) | lo;
            key = o < key ? o : key;
        }
        if ((threadIdx.x & 63u) == 0 && key != ~0ull) atomicMin(&P.key[s0], key);
    } else if (key != ~0ull) {
        atomicMin(&P.key[s], key);
    }
}

// Splits the nodes of the frontier and files their children: leaves,
// small subtrees, or nodes of the next level.  counters: [0] pool size,
// [1] next level, [2] small subtrees.
__global__ __launch_bounds__(256) void k_split(Pool P, const uint32_t* __restrict__ frontier, uint32_t nf,
                                               const GBox* __restrict__ pre, const GBox* __restrict__ suf, uint32_t L,
                                               uint32_t* counters, uint32_t* next, uint32_t* small) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const uint32_t node = frontier[f];
    const uint32_t b = P.begin[node], e = P.end[node];
    // Every node here has candidates (more than kSmallNode >= 2 * leaf + 1
    // primitives); a key no candidate wrote falls back to the first one.
    const unsigned long long key = P.key[node];
    const uint32_t mid = b + (key == ~0ull ? L : static_cast<uint32_t>(key & 0xFFFFFFFFull) + 1u);
    P.box[node] = pre[e - 1];
    P.mid[node] = mid;
    const uint32_t left = atomicAdd(&counters[0], 2u);
    P.child[node] = left;
    for (uint32_t c = 0; c < 2; ++c) {
        const uint32_t id = left + c, cb = c ? mid : b, ce = c ? e : mid, cnt = ce - cb;
        P.begin[id] = cb;
        P.end[id] = ce;
        P.box[id] = c ? suf[mid] : pre[mid - 1];
        P.key[id] = ~0ull;
        P.ic[id] = 0;
        const uint32_t kind = cnt <= L ? kLeafKind : cnt <= kSmallNode ? kSmallKind : kBigKind;
        P.kind[id] = kind;
        if (kind == kSmallKind) small[atomicAdd(&counters[2], 1u)] = id;
        if (kind == kBigKind) next[atomicAdd(&counters[1], 1u)] = id;
    }
}

__global__ __launch_bounds__(256) void k_relabel(int32_t* seg, Pool P, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t s = seg[k];
    if (s < 0) return;
    const uint32_t c = P.child[s] + (k >= P.mid[s] ? 1u : 0u);
    seg[k] = P.kind[c] == kBigKind ? static_cast<int32_t>(c) : -1;
}

__global__ __launch_bounds__(256) void k_root(Pool P, int32_t* seg, uint32_t n, uint32_t* counters, uint32_t* frontier,
                                              uint32_t* small, uint32_t small_max) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool big = n > small_max;
    if (k == 0) {
        P.begin[0] = 0;
        P.end[0] = n;
        P.key[0] = ~0ull;
        P.ic[0] = 0;
        P.rank[0] = 0;
        P.id[0] = 0;
        P.kind[0] = big ? kBigKind : kSmallKind;
        counters[0] = 1;
        counters[1] = big ? 1u : 0u;
        counters[2] = big ? 0u : 1u;
        counters[3] = 0;
        if (big) frontier[0] = 0;
        else small[0] = 0;
    }
    if (k < n) seg[k] = big ? 0 : -1;
}

__global__ __launch_bounds__(256) void k_nan_check(const float* boxes, size_t count, uint32_t* flag) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count && boxes[i] != boxes[i]) atomicOr(flag, 1u);
}

// getSplitIndex_SAH (BVH.hpp:274-311) as the host restatement computes it.
__device__ uint32_t split_sah_seq(const float* boxes, uint32_t base, uint32_t N, uint32_t L, float* left,
                                  float* right) {
    for (uint32_t i = 0; i < N - L; ++i) {
        left[i] = 0.0f;
        right[i] = 0.0f;
    }
    Box lb = to_box(load_box(boxes, base));
    left[0] = surface_area(lb);
    for (uint32_t i = 1; i < N - L; ++i) {
        lb = surrounding(lb, to_box(load_box(boxes, base + i)));
        left[i] = surface_area(lb);
    }
    Box rb = to_box(load_box(boxes, base + N - 1));
    right[N - L - 1] = surface_area(rb);
    for (uint32_t i = N - 2; i > L; --i) {
        rb = surrounding(rb, to_box(load_box(boxes, base + i)));
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

// One thread per small subtree: the reference builder over its range with
// local ids (root 0, children of the local interior node of pre-order rank
// k at 2k + 1 and 2k + 2), into lnodes[2 * begin ...].
__global__ __launch_bounds__(64) void k_small(Pool P, const uint32_t* __restrict__ small, uint32_t ns,
                                              const float* __restrict__ boxes, RefNode* lnodes, float* scratch,
                                              uint32_t L) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns) return;
    const uint32_t node = small[j];
    const uint32_t B = P.begin[node], n = P.end[node] - B;
    RefNode* out = lnodes + 2 * static_cast<size_t>(B);
    float* left = scratch + 2 * static_cast<size_t>(B);
    float* right = left + n;
    uint32_t sId[kSmallNode], sBegin[kSmallNode], sEnd[kSmallNode];
    uint32_t sp = 0, id = 0, begin = 0, end = n, maxId = 0;
    while (true) {
        Box bx = to_box(load_box(boxes, B + begin));
        for (uint32_t i = begin + 1; i < end; ++i) bx = surrounding(to_box(load_box(boxes, B + i)), bx);
        RefNode nd;
        nd.box = bx;
        const uint32_t cnt = end - begin;
        if (cnt <= L) {
            nd.offset = B + begin;
            nd.count = cnt;
            out[id] = nd;
            if (sp == 0) break;
            --sp;
            id = sId[sp];
            begin = sBegin[sp];
            end = sEnd[sp];
        } else {
            const uint32_t l = maxId + 1, r = l + 1;
            nd.offset = l;
            nd.count = 0;
            out[id] = nd;
            maxId = r > maxId ? r : maxId;
            const uint32_t split = cnt <= 2 * L ? L : split_sah_seq(boxes, B + begin, cnt, L, left, right);
            sId[sp] = r;
            sBegin[sp] = begin + split;
            sEnd[sp] = end;
            ++sp;
            id = l;
            end = begin + split;
        }
    }
    P.ic[node] = maxId / 2u;
}

// Interior-node counts, bottom up over one level's pool range.
__global__ __launch_bounds__(256) void k_count(Pool P, uint32_t lo, uint32_t hi) {
    const uint32_t p = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= hi) return;
    if (P.kind[p] == kBigKind) P.ic[p] = 1u + P.ic[P.child[p]] + P.ic[P.child[p] + 1u];
    else if (P.kind[p] == kLeafKind) P.ic[p] = 0;
}

// Pre-order ranks and ids of the children of one level's split nodes.
__global__ __launch_bounds__(256) void k_number(Pool P, uint32_t lo, uint32_t hi) {
    const uint32_t p = lo + blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= hi || P.kind[p] != kBigKind) return;
    const uint32_t r = P.rank[p], c = P.child[p];
    P.id[c] = 2u * r + 1u;
    P.id[c + 1u] = 2u * r + 2u;
    P.rank[c] = r + 1u;
    P.rank[c + 1u] = r + 1u + P.ic[c];
}

// Ids stay below 2n - 1 by construction; cap guards the output anyway.
__global__ __launch_bounds__(256) void k_emit_pool(Pool P, uint32_t np, RefNode* out, uint32_t cap) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np || P.kind[p] == kSmallKind || P.id[p] >= cap) return;
    RefNode nd;
    nd.box = to_box(P.box[p]);
    if (P.kind[p] == kLeafKind) {
        nd.offset = P.begin[p];
        nd.count = P.end[p] - P.begin[p];
    } else {
        nd.offset = 2u * P.rank[p] + 1u;
        nd.count = 0;
    }
    out[P.id[p]] = nd;
}

__global__ __launch_bounds__(64) void k_emit_small(Pool P, const uint32_t* __restrict__ small, uint32_t ns,
                                                   const RefNode* __restrict__ lnodes, RefNode* out, uint32_t cap) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ns) return;
    const uint32_t node = small[j];
    const uint32_t B = P.begin[node], R = P.rank[node], total = 2u * P.ic[node] + 1u;
    const RefNode* src = lnodes + 2 * static_cast<size_t>(B);
    for (uint32_t k = 0; k < total; ++k) {
        RefNode nd = src[k];
        if (nd.count == 0) nd.offset += 2u * R;
        const uint32_t gid = k == 0 ? P.id[node] : k + 2u * R;
        if (gid < cap) out[gid] = nd;
    }
}

uint32_t blocks(size_t n, uint32_t b) { return static_cast<uint32_t>((n + b - 1) / b); }

// Sub-allocation of the scratch buffer: base == nullptr only sizes.
struct Carver {
    char* base;
    size_t off = 0;
    template <class T>
    T* take(size_t count) {
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += (count * sizeof(T) + 255) / 256 * 256;
        return p;
    }
};

template <bool REV>
hipError_t scan(ScanBuf s, hipStream_t st) {
    const uint32_t nb = blocks(s.n, kScanBlock);
    k_scan_local<REV><<<nb, kScanThreads, 0, st>>>(s);
    k_scan_blocks<REV><<<1, 256, 0, st>>>(s, nb);
    k_scan_apply<REV><<<nb, kScanThreads, 0, st>>>(s);
    return hipGetLastError();
}

}  // namespace

#define BVH_HIP(expr)                         \
    do {                                      \
        const hipError_t e_ = (expr);         \
        if (e_ != hipSuccess) return e_;      \
    } while (0)

hipError_t gpu_build_bvh(const float* boxes6, uint32_t n, uint32_t leaf, RefNode* nodes, uint32_t* count,
                         BvhGpuScratch* sc, hipStream_t st) {
    if (leaf < 1 || leaf > 16) return hipErrorInvalidValue;
    if (n == 0) {
        const RefNode empty{Box{f3{0, 0, 0}, f3{0, 0, 0}}, 0, 0};
        BVH_HIP(hipMemcpyAsync(nodes, &empty, sizeof(empty), hipMemcpyHostToDevice, st));
        BVH_HIP(hipStreamSynchronize(st));
        *count = 1;
        return hipSuccess;
    }
    if (n > (1u << 29)) return hipErrorNotSupported;
    const size_t N = n, pool = 2 * N + 1, nb = blocks(N, kScanBlock);
    // Scratch layout (256-byte aligned pieces); a first pass only sizes it.
    struct Bufs {
        uint32_t* counters;
        int32_t* seg;
        GBox *pre, *suf;
        Pool P;
        uint32_t *frontier, *next, *small;
        RefNode* lnodes;
        float* fscratch;
        ScanBuf sbuf;
    };
    auto layout = [&](char* base, Bufs* o) {
        Carver c{base};
        o->counters = c.take<uint32_t>(8);
        o->seg = c.take<int32_t>(N);
        o->pre = c.take<GBox>(N);
        o->suf = c.take<GBox>(N);
        o->P.begin = c.take<uint32_t>(pool);
        o->P.end = c.take<uint32_t>(pool);
        o->P.box = c.take<GBox>(pool);
        o->P.child = c.take<uint32_t>(pool);
        o->P.kind = c.take<uint32_t>(pool);
        o->P.key = c.take<unsigned long long>(pool);
        o->P.mid = c.take<uint32_t>(pool);
        o->P.ic = c.take<uint32_t>(pool);
        o->P.rank = c.take<uint32_t>(pool);
        o->P.id = c.take<uint32_t>(pool);
        o->frontier = c.take<uint32_t>(N);
        o->next = c.take<uint32_t>(N);
        o->small = c.take<uint32_t>(N);
        o->lnodes = c.take<RefNode>(2 * N);
        o->fscratch = c.take<float>(2 * N);
        o->sbuf = ScanBuf{boxes6, o->seg, nullptr, n, c.take<GBox>(nb), c.take<uint32_t>(nb), c.take<uint32_t>(nb),
                          c.take<GBox>(nb)};
        return c.off;
    };
    Bufs q{};
    const size_t need = layout(nullptr, &q);
    if (sc->bytes < need) {
        if (sc->p) BVH_HIP(hipFree(sc->p));
        sc->p = nullptr;
        sc->bytes = 0;
        BVH_HIP(hipMalloc(&sc->p, need));
        sc->bytes = need;
    }
    layout(static_cast<char*>(sc->p), &q);
    uint32_t* counters = q.counters;
    int32_t* seg = q.seg;
    GBox* pre = q.pre;
    GBox* suf = q.suf;
    Pool P = q.P;
    uint32_t* frontier = q.frontier;
    uint32_t* next = q.next;
    uint32_t* small = q.small;
    RefNode* lnodes = q.lnodes;
    float* fscratch = q.fscratch;
    ScanBuf sbuf = q.sbuf;

    k_root<<<blocks(N, 256), 256, 0, st>>>(P, seg, n, counters, frontier, small, kSmallNode);
    k_nan_check<<<blocks(6 * N, 256), 256, 0, st>>>(boxes6, 6 * N, counters + 3);
    BVH_HIP(hipGetLastError());
    uint32_t host[4];
    BVH_HIP(hipMemcpyAsync(host, counters, sizeof(host), hipMemcpyDeviceToHost, st));
    BVH_HIP(hipStreamSynchronize(st));
    if (host[3]) return hipErrorNotSupported;
    // Pool index range of every level (children are allocated after their
    // parents, level by level).
    std::vector<uint32_t> level_start{0u, 1u};
    uint32_t nf = host[1];
    while (nf > 0) {
        if (level_start.size() > kMaxLevels) return hipErrorNotSupported;
        sbuf.out = pre;
        BVH_HIP(scan<false>(sbuf, st));
        sbuf.out = suf;
        BVH_HIP(scan<true>(sbuf, st));
        k_sah<<<blocks(N, 256), 256, 0, st>>>(seg, pre, suf, P, n, leaf);
        BVH_HIP(hipMemsetAsync(counters + 1, 0, sizeof(uint32_t), st));
        k_split<<<blocks(nf, 256), 256, 0, st>>>(P, frontier, nf, pre, suf, leaf, counters, next, small);
        k_relabel<<<blocks(N, 256), 256, 0, st>>>(seg, P, n);
        BVH_HIP(hipGetLastError());
        BVH_HIP(hipMemcpyAsync(host, counters, sizeof(host), hipMemcpyDeviceToHost, st));
        BVH_HIP(hipStreamSynchronize(st));
        level_start.push_back(host[0]);
        nf = host[1];
        std::swap(frontier, next);
    }
    const uint32_t npool = host[0], ns = host[2];
    if (ns) k_small<<<blocks(ns, 64), 64, 0, st>>>(P, small, ns, boxes6, lnodes, fscratch, leaf);
    // level_start[d] .. level_start[d + 1]: the nodes of level d.
    const size_t levels = level_start.size() - 1;
    for (size_t d = levels; d-- > 0;) {
        const uint32_t lo = level_start[d], hi = level_start[d + 1];
        if (hi > lo) k_count<<<blocks(hi - lo, 256), 256, 0, st>>>(P, lo, hi);
    }
    for (size_t d = 0; d < levels; ++d) {
        const uint32_t lo = level_start[d], hi = level_start[d + 1];
        if (hi > lo) k_number<<<blocks(hi - lo, 256), 256, 0, st>>>(P, lo, hi);
    }
    const uint32_t cap = 2u * n - 1u;
    k_emit_pool<<<blocks(npool, 256), 256, 0, st>>>(P, npool, nodes, cap);
    if (ns) k_emit_small<<<blocks(ns, 64), 64, 0, st>>>(P, small, ns, lnodes, nodes, cap);
    BVH_HIP(hipGetLastError());
    uint32_t ic0 = 0;
    BVH_HIP(hipMemcpyAsync(&ic0, P.ic, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    BVH_HIP(hipStreamSynchronize(st));
    if (2u * ic0 + 1u > cap) return hipErrorUnknown;
    *count = 2u * ic0 + 1u;
    return hipSuccess;
}

}  // namespace mrt
