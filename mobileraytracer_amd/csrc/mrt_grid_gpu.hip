// GPU build of the RegularGrid cell lists (RegularGrid.hpp:50-154): every
// primitive registers in the cells of its box range whose box it overlaps
// (the shared intersectBox tests of mrt_grid_reg.hpp), and each cell lists
// its primitives in ascending order, as the reference's emplace_back in
// scene order leaves them.
//
// One thread per primitive counts its cells (and every cell's primitives,
// by atomics); exclusive scans give each primitive the start of its
// (cell, primitive) pairs, written in primitive order; a stable radix sort
// by cell then yields the cell lists in ascending primitive order, and the
// scan of the cell counts their offsets.
#include "mrt_grid_gpu.hpp"

#include <algorithm>
#include <hipcub/hipcub.hpp>

#include "mrt_grid_reg.hpp"

namespace mrt {

namespace {

template <int CAT>
__device__ __forceinline__ Box prim_box(const float* p, uint32_t i, TriLayout L) {
    if constexpr (CAT == MRT_CAT_TRIANGLE) {
        const float* t = p + L.stride * static_cast<size_t>(i);
        return triangle_box_of(f3{t[L.a], t[L.a + 1], t[L.a + 2]}, f3{t[L.ab], t[L.ab + 1], t[L.ab + 2]},
                               f3{t[L.ac], t[L.ac + 1], t[L.ac + 2]});
    } else if constexpr (CAT == MRT_CAT_SPHERE) {
        const float* s = p + 5 * static_cast<size_t>(i);
        return sphere_box_of(f3{s[0], s[1], s[2]}, s[3]);
    } else {
        const float* q = p + 7 * static_cast<size_t>(i);
        return plane_box_of(f3{q[0], q[1], q[2]}, f3{q[3], q[4], q[5]});
    }
}

template <int CAT>
__device__ __forceinline__ bool prim_hits(const float* p, uint32_t i, const Box& cell, TriLayout L) {
    if constexpr (CAT == MRT_CAT_TRIANGLE) {
        const float* t = p + L.stride * static_cast<size_t>(i);
        return triangle_box_overlap(f3{t[L.a], t[L.a + 1], t[L.a + 2]}, f3{t[L.ab], t[L.ab + 1], t[L.ab + 2]},
                                    f3{t[L.ac], t[L.ac + 1], t[L.ac + 2]}, cell);
    } else if constexpr (CAT == MRT_CAT_SPHERE) {
        const float* s = p + 5 * static_cast<size_t>(i);
        return sphere_box_overlap(f3{s[0], s[1], s[2]}, s[3], cell);
    } else {
        const float* q = p + 7 * static_cast<size_t>(i);
        return plane_box_overlap(f3{q[0], q[1], q[2]}, f3{q[3], q[4], q[5]}, cell);
    }
}

__device__ __forceinline__ bool finite_box(const Box& b) {
    return isfinite(b.mn.x) && isfinite(b.mn.y) && isfinite(b.mn.z) && isfinite(b.mx.x) && isfinite(b.mx.y) &&
           isfinite(b.mx.z);
}

// EMIT = false: count each primitive's cells and each cell's primitives;
// EMIT = true: write the (cell, primitive) pairs from the primitive's start.
template <int CAT, bool EMIT>
__global__ __launch_bounds__(256) void k_grid_cells(const float* __restrict__ prims, TriLayout L, uint32_t n, GridGeom g,
                                                    uint32_t* prim_count, uint32_t* cell_count,
                                                    const uint32_t* __restrict__ prim_start, uint32_t* keys,
                                                    uint32_t* vals, uint32_t* flag) {
    const uint32_t pi = blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= n) return;
    const Box b = prim_box<CAT>(prims, pi, L);
    if (!finite_box(b)) {
        if (!EMIT) {
            atomicOr(flag, 1u);
            prim_count[pi] = 0;
        }
        return;
    }
    int32_t x1, x2, y1, y2, z1, z2;
    grid_range(g, 0, b.mn.x, b.mx.x, &x1, &x2);
    grid_range(g, 1, b.mn.y, b.mx.y, &y1, &y2);
    grid_range(g, 2, b.mn.z, b.mx.z, &z1, &z2);
    // A box below the bounds on some axis (not for a scene's own bounds)
    // registers nowhere rather than outside the grid.
    if (x1 < 0 || y1 < 0 || z1 < 0) {
        if (!EMIT) prim_count[pi] = 0;
        return;
    }
    uint32_t k = EMIT ? prim_start[pi] : 0u;
    const size_t gs = static_cast<size_t>(g.gs);
    for (int32_t x = x1; x <= x2; ++x)
        for (int32_t y = y1; y <= y2; ++y)
            for (int32_t z = z1; z <= z2; ++z) {
                if (!prim_hits<CAT>(prims, pi, grid_cell(g, x, y, z), L)) continue;
                const uint32_t idx = static_cast<uint32_t>(x + y * gs + z * gs * gs);
                if (EMIT) {
                    keys[k] = idx;
                    vals[k] = pi;
                } else {
                    atomicAdd(&cell_count[idx], 1u);
                }
                ++k;
            }
    if (!EMIT) prim_count[pi] = k;
}

template <int CAT, bool EMIT>
void launch_cells(const float* prims, TriLayout L, uint32_t n, const GridGeom& g, uint32_t* prim_count,
                  uint32_t* cell_count, const uint32_t* prim_start, uint32_t* keys, uint32_t* vals, uint32_t* flag,
                  hipStream_t st) {
    k_grid_cells<CAT, EMIT><<<(n + 255) / 256, 256, 0, st>>>(prims, L, n, g, prim_count, cell_count, prim_start,
                                                              keys, vals, flag);
}

template <bool EMIT>
void cells(int category, const float* prims, TriLayout L, uint32_t n, const GridGeom& g, uint32_t* prim_count,
           uint32_t* cell_count, const uint32_t* prim_start, uint32_t* keys, uint32_t* vals, uint32_t* flag,
           hipStream_t st) {
    if (category == MRT_CAT_TRIANGLE)
        launch_cells<MRT_CAT_TRIANGLE, EMIT>(prims, L, n, g, prim_count, cell_count, prim_start, keys, vals, flag,
                                             st);
    else if (category == MRT_CAT_SPHERE)
        launch_cells<MRT_CAT_SPHERE, EMIT>(prims, L, n, g, prim_count, cell_count, prim_start, keys, vals, flag, st);
    else
        launch_cells<MRT_CAT_PLANE, EMIT>(prims, L, n, g, prim_count, cell_count, prim_start, keys, vals, flag, st);
}

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

hipError_t grow(GridGpuScratch::Buf* b, size_t bytes) {
    if (b->bytes >= bytes) return hipSuccess;
    if (b->p) {
        const hipError_t e = hipFree(b->p);
        if (e != hipSuccess) return e;
    }
    b->p = nullptr;
    b->bytes = 0;
    const hipError_t e = hipMalloc(&b->p, bytes);
    if (e == hipSuccess) b->bytes = bytes;
    return e;
}

}  // namespace

GridGpuScratch::~GridGpuScratch() {
    for (Buf* b : {&counts, &pairs})
        if (b->p) (void)hipFree(b->p);
}

#define GRID_HIP(expr)                    \
    do {                                  \
        const hipError_t e_ = (expr);     \
        if (e_ != hipSuccess) return e_;  \
    } while (0)

hipError_t gpu_build_grid(int category, const float* prims, TriLayout L, uint32_t n, const Box& bounds, int32_t gs,
                          GridGpuScratch* sc, GridGpuOut* out, hipStream_t st) {
    if (gs <= 0 || gs > 1024 || n >= (1u << 30)) return hipErrorInvalidValue;
    const GridGeom g = grid_geom(bounds, gs);
    const size_t ncell = static_cast<size_t>(gs) * gs * gs;
    if (ncell + 1 >= (1ull << 32)) return hipErrorInvalidValue;
    // Counting pieces: counters, per-primitive counts and starts, per-cell
    // counts and offsets, the scans' temporary storage.
    size_t scanTemp = 0, scanTemp2 = 0;
    GRID_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scanTemp, static_cast<uint32_t*>(nullptr),
                                              static_cast<uint32_t*>(nullptr), static_cast<int>(n + 1), st));
    GRID_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scanTemp2, static_cast<uint32_t*>(nullptr),
                                              static_cast<uint32_t*>(nullptr), static_cast<int>(ncell + 1), st));
    scanTemp = std::max(scanTemp, scanTemp2);
    auto layout = [&](char* base, uint32_t** flag, uint32_t** pc, uint32_t** ps, uint32_t** cc, uint32_t** co,
                      void** tmp) {
        Carve c{base};
        *flag = c.take<uint32_t>(4);
        *pc = c.take<uint32_t>(n + 1);
        *ps = c.take<uint32_t>(n + 1);
        *cc = c.take<uint32_t>(ncell + 1);
        *co = c.take<uint32_t>(ncell + 1);
        *tmp = c.take<char>(scanTemp);
        return c.off;
    };
    uint32_t *flag, *pc, *ps, *cc, *co;
    void* tmp;
    GRID_HIP(grow(&sc->counts, layout(nullptr, &flag, &pc, &ps, &cc, &co, &tmp)));
    layout(static_cast<char*>(sc->counts.p), &flag, &pc, &ps, &cc, &co, &tmp);
    GRID_HIP(hipMemsetAsync(flag, 0, 4 * sizeof(uint32_t), st));
    GRID_HIP(hipMemsetAsync(pc + n, 0, sizeof(uint32_t), st));
    GRID_HIP(hipMemsetAsync(cc, 0, (ncell + 1) * sizeof(uint32_t), st));
    if (n) cells<false>(category, prims, L, n, g, pc, cc, nullptr, nullptr, nullptr, flag, st);
    GRID_HIP(hipGetLastError());
    size_t t = scanTemp;
    GRID_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, t, pc, ps, static_cast<int>(n + 1), st));
    t = scanTemp;
    GRID_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, t, cc, co, static_cast<int>(ncell + 1), st));
    uint32_t host[2] = {0, 0};
    GRID_HIP(hipMemcpyAsync(&host[0], flag, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    GRID_HIP(hipMemcpyAsync(&host[1], ps + n, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    GRID_HIP(hipStreamSynchronize(st));
    if (host[0]) return hipErrorNotSupported;
    const uint32_t total = host[1];
    // Pairs: keys and values, their sorted copies, the sort's storage.
    int endBit = 1;
    while ((1ull << endBit) < ncell) ++endBit;
    size_t sortTemp = 0;
    GRID_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sortTemp, static_cast<uint32_t*>(nullptr),
                                                static_cast<uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
                                                static_cast<uint32_t*>(nullptr), static_cast<int>(total), 0, endBit,
                                                st));
    auto playout = [&](char* base, uint32_t** k, uint32_t** v, uint32_t** k2, uint32_t** v2, void** tmp2) {
        Carve c{base};
        *k = c.take<uint32_t>(total + 1);
        *v = c.take<uint32_t>(total + 1);
        *k2 = c.take<uint32_t>(total + 1);
        *v2 = c.take<uint32_t>(total + 1);
        *tmp2 = c.take<char>(sortTemp);
        return c.off;
    };
    uint32_t *keys, *vals, *keys2, *vals2;
    void* tmp2;
    GRID_HIP(grow(&sc->pairs, playout(nullptr, &keys, &vals, &keys2, &vals2, &tmp2)));
    playout(static_cast<char*>(sc->pairs.p), &keys, &vals, &keys2, &vals2, &tmp2);
    if (total) {
        cells<true>(category, prims, L, n, g, nullptr, nullptr, ps, keys, vals, flag, st);
        GRID_HIP(hipGetLastError());
        t = sortTemp;
        GRID_HIP(hipcub::DeviceRadixSort::SortPairs(tmp2, t, keys, keys2, vals, vals2, static_cast<int>(total), 0,
                                                    endBit, st));
    }
    GRID_HIP(hipStreamSynchronize(st));
    out->offsets = co;
    out->refs = vals2;
    out->n_refs = total;
    return hipSuccess;
}

}  // namespace mrt
