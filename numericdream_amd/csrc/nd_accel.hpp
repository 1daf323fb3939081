// nd_accel.hpp - the empty-space accelerator on the device: cascaded grids of safe skip
// distances around the camera (cs_build_accel, template.wgsl:5409-5431) and their lookup
// (accel_skip, :426-447).
#pragma once

#include "nd_de.hpp"

namespace nd {

// What a ray needs to look the grid up, held in registers: the grid, the cascade table and
// the centre and half-extent of the finest cascade.
struct AccelView {
    const float* grid;  // null: accelerator off
    const float* prm;   // [level][corner xyz, cell]
    float cx, cy, cz;   // centre of level 0
    float inv_half0;    // 1 / half-extent of level 0
    int res, levels;
};

NDD AccelView accel_view(const DevJob& J) {
    AccelView a;
    a.grid = J.accel;
    a.prm = &J.accel_prm[0][0];
    a.res = J.accel_res;
    a.levels = J.accel_levels;
    const float half0 = 0.5f * (float)J.accel_res * J.accel_prm[0][3];
    a.cx = J.accel_prm[0][0] + half0;
    a.cy = J.accel_prm[0][1] + half0;
    a.cz = J.accel_prm[0][2] + half0;
    a.inv_half0 = 1.f / half0;
    return a;
}

// Safe skip distance at p: the grid value of the finest level whose grid holds p, 0 outside
// every grid (accel_skip). Level l is centred within one of its cells of level 0's centre
// and reaches half0 2^l (1 + 2/res) from it at most, so no level below
// ceil(log2(|p - centre|_inf / half0)) - 1 can hold p; the search starts one level lower
// still (the hardware log2 is accurate to an ulp), which finds the same level as a search
// from level 0. The log is clamped at 0 before the conversion: at the centre itself (u = 0)
// it is -inf, whose conversion saturates to INT_MIN and would wrap past every level.
// Cells are indexed by floor((p - corner) / cell), the quotient rounded as a correctly rounded
// division rounds it (cell_index).
NDD float cell_index(float x, float cell, float ic) {
    // The product with the hardware reciprocal is corrected once by its exact residual
    // (fma(-q, cell, x)), which lands exact quotients on their integer: the camera sits exactly
    // on a cell boundary of every cascade (their common centre), so every camera ray's first
    // lookup is such a quotient, and the uncorrected product can fall just below it.
    const float q0 = x * ic;
    return floorf(fmaf(fmaf(-q0, cell, x), ic, q0));
}
// The cell that holds p: its level and coordinates; false outside every grid.
NDD bool accel_locate(const AccelView& A, const V3<float>& p, int* lvl, int* ix, int* iy, int* iz) {
    const float u = fmaxf(fabsf(p.x - A.cx), fmaxf(fabsf(p.y - A.cy), fabsf(p.z - A.cz)));
    int l = (int)ceilf(fmaxf(__builtin_amdgcn_logf(u * A.inv_half0), 0.f)) - 2;
    l = l < 0 ? 0 : l;
    const float fres = (float)A.res;
    for (int k = 0; k < 4 && l < A.levels; ++k, ++l) {
        const float* q = A.prm + 4 * l;
        const float ic = __builtin_amdgcn_rcpf(q[3]);
        const float fx = cell_index(p.x - q[0], q[3], ic);
        const float fy = cell_index(p.y - q[1], q[3], ic);
        const float fz = cell_index(p.z - q[2], q[3], ic);
        // written so that a NaN coordinate (a ray whose DE overflowed) is outside every grid
        if (!(fx >= 0.f && fy >= 0.f && fz >= 0.f && fx < fres && fy < fres && fz < fres)) continue;
        *lvl = l; *ix = (int)fx; *iy = (int)fy; *iz = (int)fz;
        return true;
    }
    return false;
}
// Index into the grid ([level][z][y][x]) of the cell that holds p, -1 outside every grid. A
// grid of one scene holds fewer than 2^31 cells (checked where it is built).
NDD int accel_index(const AccelView& A, const V3<float>& p) {
    int l, ix, iy, iz;
    if (!accel_locate(A, p, &l, &ix, &iy, &iz)) return -1;
    return ((l * A.res + iz) * A.res + iy) * A.res + ix;
}
NDD float accel_skip(const AccelView& A, const V3<float>& p) {
    const int idx = accel_index(A, p);
    return idx < 0 ? 0.f : A.grid[idx];
}

// The last cell a ray read and its value. Consecutive probes of a ray near a surface step by
// less than a cell, so most land in the cell of the previous probe: the value is then reused
// from registers instead of gathered again. Reset whenever the lane takes a ray whose job may
// have another grid.
struct AccelCache {
    int idx;  // -2: empty
    float val;
};
NDD void accel_cache_reset(AccelCache& c) { c.idx = -2; c.val = 0.f; }
NDD float accel_skip_cached(const AccelView& A, AccelCache& c, const V3<float>& p) {
    const int idx = accel_index(A, p);
    if (idx != c.idx) {
        c.val = idx < 0 ? 0.f : A.grid[idx];
        c.idx = idx;
    }
    return c.val;
}

// The accelerator test of a shadow-ray step (visibility_march, template.wgsl:2297-2303): a
// skip is taken where it cannot lower the penumbra estimate (skip > SHADOW_MIN_STEP and
// k skip >= res t); it uses up one of the ray's steps.
NDD bool shadow_skip_applies(float skip, float k, float res, float t) {
    return skip > 0.001f && k * skip >= res * t;
}

}  // namespace nd
