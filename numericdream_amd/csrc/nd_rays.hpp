// nd_rays.hpp - primary rays and ray-slot bookkeeping shared by the HIP passes.
#pragma once

#include "nd_de.hpp"

namespace nd {

// ----------------------------------------------------------------------------- rays
// Primary ray (cs_wf_generate, template.wgsl:5119-5127).
NDD V3<float> primary_dir(const DevJob& J, int px, int py) {
    float nx = ((float)px + 0.5f) * J.inv_tw * 2.f - 1.f;
    float ny = 1.f - ((float)py + 0.5f) * J.inv_th * 2.f;
    float pfx = nx * J.tile_scale[0] + J.tile_bias[0];
    float pfy = ny * J.tile_scale[1] + J.tile_bias[1];
    float ux = pfx * J.aspect, uy = -pfy;
    float dx = J.fwd[0] + ux * J.right[0] + uy * J.up[0];
    float dy = J.fwd[1] + ux * J.right[1] + uy * J.up[1];
    float dz = J.fwd[2] + ux * J.right[2] + uy * J.up[2];
    float il = __frsqrt_rn(dx * dx + dy * dy + dz * dz);
    return mk3(dx * il, dy * il, dz * il);
}

// Ray slot -> (job, x, y inside the rect). Slots enumerate the tiles of each job, 64 per tile
// (2^tile_lg wide, 64 / 2^tile_lg high), row by row within a tile.
NDD int find_job(const uint32_t* __restrict__ job_base, int n_jobs, uint32_t r) {
    int lo = 0, hi = n_jobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (job_base[mid] <= r) lo = mid; else hi = mid - 1;
    }
    return lo;
}
NDD bool slot_pixel(const DevJob& J, uint32_t local, int* x, int* y) {
    uint32_t tile = local >> 6, l = local & 63u;
    int tx = (int)(tile % (uint32_t)J.tiles_x), ty = (int)(tile / (uint32_t)J.tiles_x);
    *x = (tx << J.tile_lg) + (int)(l & ((1u << J.tile_lg) - 1u));
    *y = ty * (64 >> J.tile_lg) + (int)(l >> J.tile_lg);
    return *x < J.rect_w && *y < J.rect_h;
}

}  // namespace nd
