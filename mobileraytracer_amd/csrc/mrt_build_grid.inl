// RegularGrid build (template part of mrt_build.hpp).
#pragma once
#include <algorithm>
#include <cmath>

namespace mrt {

template <class BoxFn, class HitFn>
void build_grid(uint32_t n, const Box& bounds, int32_t gs, BoxFn box_of, HitFn intersects, GridOut* out) {
    out->grid_size = gs;
    uint32_t bits = 0;
    for (uint32_t v = static_cast<uint32_t>(gs); v > 0; v >>= 1) ++bits;
    out->grid_shift = bits - 1u;
    out->ext = bounds;
    const f3 e = sub(bounds.mx, bounds.mn);
    const float g = static_cast<float>(gs);
    out->sr = f3{g / e.x, g / e.y, g / e.z};
    out->cw = muls(e, 1.0f / g);
    const size_t ncell = static_cast<size_t>(gs) * gs * gs;
    const float sizeX = bounds.mx.x - bounds.mn.x, sizeY = bounds.mx.y - bounds.mn.y, sizeZ = bounds.mx.z - bounds.mn.z;
    const float dx = sizeX / g, dy = sizeY / g, dz = sizeZ / g;
    const float dxr = dx > 0.0f ? 1.0f / dx : 1.0f;
    const float dyr = dy > 0.0f ? 1.0f / dy : 1.0f;
    const float dzr = dz > 0.0f ? 1.0f / dz : 1.0f;
    // Two passes (count, fill) over the same candidate cells so the CSR lists
    // come out in ascending primitive order, as the reference's
    // emplace_back in scene order produces.
    std::vector<uint32_t> counts(ncell + 1, 0u);
    std::vector<std::pair<uint32_t, uint32_t>> pairs;  // (cell, prim)
    pairs.reserve(static_cast<size_t>(n) * 4);
    for (uint32_t pi = 0; pi < n; ++pi) {
        const Box b = box_of(pi);
        auto range = [&](float lo, float hi, float mn, float rec, float size, int32_t* a, int32_t* c) {
            int32_t x1 = static_cast<int32_t>((lo - mn) * rec);
            int32_t x2 = static_cast<int32_t>((hi - mn) * rec) + 1;
            x1 = std::max(0, x1);
            x2 = std::min(gs - 1, x2);
            x2 = std::fabs(size) < kEpsilon ? 0 : x2;
            x1 = std::min(x1, x2);
            *a = x1;
            *c = x2;
        };
        int32_t x1, x2, y1, y2, z1, z2;
        range(b.mn.x, b.mx.x, bounds.mn.x, dxr, sizeX, &x1, &x2);
        range(b.mn.y, b.mx.y, bounds.mn.y, dyr, sizeY, &y1, &y2);
        range(b.mn.z, b.mx.z, bounds.mn.z, dzr, sizeZ, &z1, &z2);
        for (int32_t x = x1; x <= x2; ++x)
            for (int32_t y = y1; y <= y2; ++y)
                for (int32_t z = z1; z <= z2; ++z) {
                    const size_t idx = static_cast<size_t>(x) + static_cast<size_t>(y) * gs +
                                       static_cast<size_t>(z) * gs * gs;
                    const f3 pos{bounds.mn.x + static_cast<float>(x) * dx, bounds.mn.y + static_cast<float>(y) * dy,
                                 bounds.mn.z + static_cast<float>(z) * dz};
                    const Box cell{pos, add(pos, f3{dx, dy, dz})};
                    if (intersects(pi, cell)) {
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
