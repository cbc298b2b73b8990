// RegularGrid build (template part of mrt_build.hpp).
#pragma once
#include <algorithm>
#include <cmath>

#include "mrt_grid_reg.hpp"

namespace mrt {

// Size, bounds and walk constants of a grid (the cell lists aside).
inline void grid_meta(const Box& bounds, int32_t gs, GridOut* out) {
    out->grid_size = gs;
    uint32_t bits = 0;
    for (uint32_t v = static_cast<uint32_t>(gs); v > 0; v >>= 1) ++bits;
    out->grid_shift = bits - 1u;
    out->ext = bounds;
    const f3 e = sub(bounds.mx, bounds.mn);
    const float g = static_cast<float>(gs);
    out->sr = f3{g / e.x, g / e.y, g / e.z};
    out->cw = muls(e, 1.0f / g);
}

template <class BoxFn, class HitFn>
void build_grid(uint32_t n, const Box& bounds, int32_t gs, BoxFn box_of, HitFn intersects, GridOut* out) {
    grid_meta(bounds, gs, out);
    const size_t ncell = static_cast<size_t>(gs) * gs * gs;
    const GridGeom geo = grid_geom(bounds, gs);
    // Two passes (count, fill) over the same candidate cells so the CSR lists
    // come out in ascending primitive order, as the reference's
    // emplace_back in scene order produces.
    std::vector<uint32_t> counts(ncell + 1, 0u);
    std::vector<std::pair<uint32_t, uint32_t>> pairs;  // (cell, prim)
    pairs.reserve(static_cast<size_t>(n) * 4);
    for (uint32_t pi = 0; pi < n; ++pi) {
        const Box b = box_of(pi);
        int32_t x1, x2, y1, y2, z1, z2;
        grid_range(geo, 0, b.mn.x, b.mx.x, &x1, &x2);
        grid_range(geo, 1, b.mn.y, b.mx.y, &y1, &y2);
        grid_range(geo, 2, b.mn.z, b.mx.z, &z1, &z2);
        for (int32_t x = x1; x <= x2; ++x)
            for (int32_t y = y1; y <= y2; ++y)
                for (int32_t z = z1; z <= z2; ++z) {
                    const size_t idx = static_cast<size_t>(x) + static_cast<size_t>(y) * gs +
                                       static_cast<size_t>(z) * gs * gs;
                    if (intersects(pi, grid_cell(geo, x, y, z))) {
                        pairs.emplace_back(static_cast<uint32_t>(idx), pi);
                        ++counts[idx + 1];
                    }
                }
    }
    for (size_t i = 0; i < ncell; ++i) counts[i + 1] += counts[i];
    out->offsets = counts;
    out->refs.assign(pairs.size(), 0u);
    std::vector<uint32_t> fill(counts.begin(), counts.end() - 1);
    for (const auto& pr : pairs) out->refs[fill[pr.first]++] = pr.second;
}

}  // namespace mrt
