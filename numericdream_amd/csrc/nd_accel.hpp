// nd_accel.hpp - the empty-space accelerator on the device: cascaded grids of safe skip
// distances around the camera (cs_build_accel, template.wgsl:5409-5431) and their lookup
// (accel_skip, :426-447).
#pragma once

#include "nd_de.hpp"

namespace nd {

// Safe skip distance at p: the first level, from the finest, whose grid holds p; 0 outside
// every grid. Coordinates are compared as floats before the integer conversion.
NDD float accel_skip(const DevJob& J, const V3<float>& p) {
    const int res = J.accel_res;
    const float fres = (float)res;
    for (int l = 0; l < J.accel_levels; ++l) {
        const float cell = J.accel_prm[l][3];
        if (!(cell > 0.f)) continue;
        const float fx = floorf((p.x - J.accel_prm[l][0]) / cell);
        const float fy = floorf((p.y - J.accel_prm[l][1]) / cell);
        const float fz = floorf((p.z - J.accel_prm[l][2]) / cell);
        if (fx < 0.f || fy < 0.f || fz < 0.f || fx >= fres || fy >= fres || fz >= fres) continue;
        const size_t idx = (((size_t)l * res + (size_t)fz) * res + (size_t)fy) * res + (size_t)fx;
        return J.accel[idx];
    }
    return 0.f;
}

// Accelerator steps of a shadow ray before its next DE sample (visibility_march,
// template.wgsl:2297-2303): a skip is taken where it cannot lower the penumbra estimate
// (skip > SHADOW_MIN_STEP and k skip >= res t) and uses up one of the ray's steps. Returns
// true when the ray ended (step budget or light distance reached).
NDD bool shadow_skips(const DevJob& J, const V3<float>& o, const V3<float>& d, float k, float res, float ldist,
                      int max_steps, float* t, int* step) {
    for (;;) {
        const float skip = accel_skip(J, mk3(o.x + d.x * *t, o.y + d.y * *t, o.z + d.z * *t));
        if (!(skip > 0.001f && k * skip >= res * *t)) return false;
        *t += skip;
        ++*step;
        if (*step >= max_steps || *t >= ldist) return true;
    }
}

}  // namespace nd
