// nd_api.cpp - the C ABI (include/nd_hip.h): context, error sink, Uniforms decoding into
// device job records, scheduling of the plain and persistent geometry passes, the
// autotuner and device statistics.
//
// Reference counterparts: Context.init / FractalRenderer.init (src/gpu/webgpu_context.zig:
// 243-283, src/gpu/fractal_renderer.zig:1769-), the error sink (webgpu_context.zig:131-146),
// renderToImage / renderTile / drawSlot (fractal_renderer.zig:3548-3940, 4284-4298) and
// stampFormulaCodes (fractal_renderer.zig:2882-2896).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "nd_config.h"
#include "nd_hip.h"
#include "nd_internal.h"
#include "nd_ctx.h"

// ----------------------------------------------------------------------------- errors
nd_status nd_fail(nd_ctx* c, nd_status st, const char* fmt, ...) {
    if (!c) return st;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof(c->err), fmt, ap);
    va_end(ap);
    return st;
}

nd_status nd_hip_fail(nd_ctx* c, hipError_t e, const char* what) {
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return nd_fail(c, ND_E_OOM, "%s: %s", what, hipGetErrorString(e));
    return nd_fail(c, ND_E_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

nd_status DevBuf::ensure(nd_ctx* c, size_t bytes) {
    if (bytes <= cap) return ND_OK;
    // captured graphs hold the old pointers: no graph outlives a buffer it names
    nd_drop_graphs(c);
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max(bytes, (size_t)256);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        p = nullptr;
        return nd_hip_fail(c, e, "hipMalloc");
    }
    cap = want;
    return ND_OK;
}

void nd_drop_graphs(nd_ctx* c) {
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->mc_graph_exec) (void)hipGraphExecDestroy(c->mc_graph_exec);
    c->graph_exec = nullptr;
    c->mc_graph_exec = nullptr;
}
void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

// ----------------------------------------------------------------------------- packing
static int i32(float x) { return (int)x; }

// The @param slots each formula file reads outside its step arithmetic: its de_iterations
// count (-1: one iteration), the degrees of its rotate_xyz_deg (-1: a literal 0, all -1: no
// rotation) and of its twist about z (-1: none). From the @param lines and de_iterations
// of each formula file.
struct FormulaSlots { int8_t iters, rx, ry, rz, twist; };
static const FormulaSlots kSlots[ND_FORMULA_COUNT] = {
    {1, -1, -1, -1, -1},   // mandelbrot.wgsl:3-5
    {4, 5, 6, 7, -1},      // amazingbox_box.wgsl:13-15, 29
    {0, 5, 6, 7, -1},      // menger3_ifs.wgsl:9-11, 17
    {6, -1, -1, -1, -1},   // kleinian_kleinian.wgsl:12-14
    {-1, -1, -1, -1, -1},  // sphere_primitive.wgsl:2-4
    {-1, -1, -1, -1, -1},  // cube_primitive.wgsl:2-4
    {-1, -1, -1, -1, -1},  // cylinder_primitive.wgsl:2-4
    {-1, -1, -1, -1, -1},  // torus_primitive.wgsl:3-5
    {-1, -1, -1, -1, -1},  // pyramid_primitive.wgsl:5-7
    {-1, -1, -1, -1, -1},  // prism_primitive.wgsl:4-6
    {1, 5, 6, 7, -1},      // sierpinski3_ifs.wgsl:9-11, 17
    {4, 5, 6, 7, -1},      // cubejulia_ifs_julia.wgsl:9-11, 17
    {4, 5, 6, 7, -1},      // juliakifs_ifs_julia.wgsl:9-11, 17
    {1, 5, 6, 7, -1},      // diagonaltetra_ifs_fold.wgsl:9-11, 25
    {2, 6, 7, 8, -1},      // starkifs_ifs_fold.wgsl:12-14, 31
    {5, 6, 7, 8, -1},      // kaleidojulia_ifs_fold_julia.wgsl:12-14, 31
    {2, -1, -1, -1, -1},   // bulbtetra_hybrid_bulb_ifs.wgsl:9-11
    {3, 5, 6, 7, -1},      // icosahedralifs_ifs.wgsl:9-11, 36
    {6, 8, 9, 10, -1},     // icosahedraljulia_ifs_julia.wgsl:12-14, 40
    {4, 6, 7, -1, 3},      // icosahedraltwist_ifs_twist.wgsl:9-11, 16, 41
    {3, 5, 6, 7, -1},      // diagonalicosa_ifs_fold.wgsl:9-11, 43
    {1, 5, 6, 7, -1},      // hybridkifs_hybrid_ifs.wgsl:9-11, 21
    {2, 6, 7, -1, 1},      // kaleidoscopicifs_ifs_twist.wgsl:9-11, 15, 28
    {7, 8, 9, 10, -1},     // juliabox_box_julia.wgsl:12-14, 29
    {4, 5, 6, 7, -1},      // amazingsurf_box_fold.wgsl:13-15, 29
    {4, 5, 6, 7, -1},      // mandelboxfoam_box_fold.wgsl:13-15, 29
    {4, 5, 6, 7, -1},      // mandelboxsort_box_fold.wgsl:13-15, 32
    {4, 5, 6, 7, -1},      // mandelboxdiag_box_fold.wgsl:13-15, 37
    {5, 6, 7, 8, -1},      // mandelboxinv_box_fold.wgsl:14-16, 38
    {5, 6, 7, -1, 4},      // mandelboxtwist_box_twist.wgsl:13-15, 22, 34
    {5, 6, 7, 8, -1},      // prismkifs_box_fold.wgsl:16-18, 44
    {3, 4, 5, 6, -1},      // buffalo_box.wgsl:12-14, 28
    {6, 7, 8, 9, -1},      // buffalojulia_box_julia.wgsl:11-13, 28
    {4, 5, 6, -1, 3},      // buffalotwist_box_twist.wgsl:12-14, 20, 33
    {3, 4, 5, 6, -1},      // apollonian_kleinian.wgsl:12-14, 28
    {6, 7, 8, 9, -1},      // apollonianjulia_kleinian_julia.wgsl:11-13, 28
    {7, -1, -1, -1, -1},   // kleinianjulia_kleinian_julia.wgsl:10-12
    {5, -1, -1, -1, -1},   // pseudokleinian_kleinian.wgsl:11-13
    {1, -1, -1, -1, -1},   // juliabulb_bulb_julia.wgsl:8-10
    {1, -1, -1, -1, -1},   // mandelbulbmirror_bulb_fold.wgsl:9-11
    {1, -1, -1, -1, 2},    // mandelbulbtwist_bulb_twist.wgsl:10-12, 16
    {1, -1, -1, -1, -1},   // mandelbulbdual_hybrid_bulb.wgsl:10-12
    {5, -1, -1, -1, -1},   // quatjulia_bulb_julia.wgsl:9-11
    {5, 8, 9, 10, -1},     // bulbbox_hybrid_bulb_box.wgsl:16-18, 44
    {7, -1, -1, -1, -1},   // mandalay_box_fold.wgsl:23-25
    {1, -1, -1, -1, -1},   // wobblekifs_ifs_twist_julia.wgsl:27-29
    {0, -1, -1, -1, -1},   // radioactive_ifs_carve_twist.wgsl:12-14
    {0, -1, -1, -1, -1},   // recursivesphere_ifs_carve_twist.wgsl:18-20
    {0, -1, -1, -1, -1},   // tree_ifs.wgsl:15-17
    {5, -1, -1, -1, -1},   // kleinianlimit_kleinian.wgsl:21-23
    {0, -1, -1, -1, -1},   // fbmsdf_primitive_carve.wgsl:29-31
    {-1, -1, -1, -1, -1},  // seascape_primitive.wgsl:22-24
    {-1, -1, -1, -1, -1},  // dragon_ifs_twist.wgsl:13-15
    {0, -1, -1, -1, -1},   // romanesco_ifs.wgsl:24-26
    {-1, -1, -1, -1, -1},  // heightmap_primitive.wgsl:10-12 (primitive_de_iterations: 1)
};

// I is DevInst or DevMixin (the same formula fields).
template <class I>
static void hoist(I& in) {
    std::memset(in.h, 0, sizeof(in.h));
    const float* p = in.p;
    switch (in.formula) {
        case ND_FORMULA_MANDELBULB:
            in.h[0] = p[0];
            in.h[1] = p[0] - 1.0f;
            in.bulb_pow8 = p[0] == 8.0f;
            break;
        case ND_FORMULA_MANDELBOX: {
            float m2 = p[1] * p[1], f2 = p[2] * p[2];
            in.h[0] = p[0];
            in.h[1] = std::fabs(p[0]);
            in.h[2] = m2;
            in.h[3] = f2;
            in.h[4] = f2 / m2;
            in.h[5] = p[3];
            in.h[6] = p[8];
            in.h[7] = p[9];
            in.h[8] = p[10];
            in.h[9] = p[11];
            in.h[10] = std::fmax(std::fabs(p[8]), 1.0f);
            break;
        }
        case ND_FORMULA_MENGER3:
            break;
        case ND_FORMULA_KLEINIAN:
            in.h[6] = 1.0f / p[5];
            in.h[7] = p[7];
            in.h[8] = p[8];
            in.h[9] = p[9];
            in.h[10] = p[10];
            in.h[11] = std::fmax(std::fabs(p[7]), 1.0f);
            break;
    }
    // formula-internal Euler rotation (degrees) and twist angle
    in.rot_on = 0;
    if (in.formula >= 0 && in.formula < ND_FORMULA_COUNT) {
        const FormulaSlots& fs = kSlots[in.formula];
        const float k = 3.14159265358979323846f / 180.0f;
        if (fs.twist >= 0) {
            in.h[14] = std::cos(p[fs.twist] * k);
            in.h[15] = std::sin(p[fs.twist] * k);
        }
        float rx = fs.rx >= 0 ? p[fs.rx] : 0.0f, ry = fs.ry >= 0 ? p[fs.ry] : 0.0f, rz = fs.rz >= 0 ? p[fs.rz] : 0.0f;
        if (!(rx == 0.0f && ry == 0.0f && rz == 0.0f)) {
            in.rot_on = 1;
            const float k = 3.14159265358979323846f / 180.0f;
            in.rot_cs[0] = std::cos(rx * k); in.rot_cs[1] = std::sin(rx * k);
            in.rot_cs[2] = std::cos(ry * k); in.rot_cs[3] = std::sin(ry * k);
            in.rot_cs[4] = std::cos(rz * k); in.rot_cs[5] = std::sin(rz * k);
        }
    }
}

bool nd_scene_ext(const DevScene& s) {
    for (int i = 0; i < s.count && i < ND_MAX_INSTANCES; ++i)
        if (!s.inst[i].hidden && (s.inst[i].formula > ND_FORMULA_KLEINIAN || s.inst[i].hybrid)) return true;
    return false;
}

static int formula_iters(int f, const float* p) {
    if (f < 0 || f >= ND_FORMULA_COUNT) return 0;
    return kSlots[f].iters >= 0 ? i32(p[kSlots[f].iters]) : 1;
}

nd_status nd_pack_scene(nd_ctx* c, const nd_uniforms& u, DevScene& s) {
    std::memset(&s, 0, sizeof(s));
    if (u.warp_count > 0.5f) return nd_fail(c, ND_E_UNSUPPORTED, "domain warps (warp_count) are not implemented");
    if (u.fft_active > 0.5f) return nd_fail(c, ND_E_UNSUPPORTED, "the FFT view (fft_active) is not implemented");
    if (((uint32_t)u.approx_flags & (1u << 8)) != 0u)
        return nd_fail(c, ND_E_UNSUPPORTED, "the distance LOD approximation is not implemented");
    s.count = std::max(i32(u.instance_count), 1);
    if (s.count > ND_MAX_INSTANCES) s.count = ND_MAX_INSTANCES;
    s.empty_de = u.max_dist * 2.0f + 1.0f;
    s.carve_count = std::min(std::max(i32(u.carve_count), 0), ND_MAX_CARVES);
    for (int i = 0; i < ND_MAX_CARVES; ++i)
        for (int k = 0; k < 4; ++k) s.carves[i][k] = u.carves[i][k];
    for (int i = 0; i < ND_MAX_INSTANCES; ++i) {
        const nd_fractal_instance& fi = u.instances[i];
        DevInst& in = s.inst[i];
        std::memset(&in, 0, sizeof(in));
        for (int k = 0; k < 16; ++k) in.dparam[k] = -1;
        in.hidden = fi.hidden > 0.5f;
        const int code = i32(fi.code);
        in.formula = (code >= 0 && code < ND_MAX_CODES) ? c->code_map[code] : -1;
        if (i < s.count && !in.hidden) {
            if (code >= 0 && in.formula < 0)
                return nd_fail(c, ND_E_UNSUPPORTED_FORMULA, "instance %d: formula code %d has no HIP distance estimator bound", i, code);
        }
        in.combine = i32(fi.combine_mode + 0.5f);
        in.blend_k = fi.blend_k;
        for (int k = 0; k < 16; ++k) in.p[k] = fi.params[k];
        in.iters = formula_iters(in.formula, in.p);
        float mag[3];
        bool ident = true;
        for (int k = 0; k < 3; ++k) {
            mag[k] = std::fmax(std::fabs(fi.scale[k]), 0.001f);
            float sgn = fi.scale[k] < 0.0f ? -mag[k] : mag[k];
            in.inv_s[k] = 1.0f / sgn;
            in.offset[k] = fi.offset[k];
            if (fi.offset[k] != 0.0f || sgn != 1.0f) ident = false;
            for (int r = 0; r < 3; ++r) {
                in.frame[3 * k + r] = fi.frame[k][r];
                if (fi.frame[k][r] != (k == r ? 1.0f : 0.0f)) ident = false;
            }
        }
        in.identity = ident;
        in.de_factor = std::fmin(mag[0], std::fmin(mag[1], mag[2])) * std::fmin(std::fmax(fi.step_safety, 0.001f), 1.0f);
        in.trap_shape = i32(fi.trap_shape + 0.5f);
        in.trap_mode = i32(fi.trap_mode + 0.5f);
        for (int k = 0; k < 3; ++k) {
            in.trap_center[k] = fi.trap_center[k];
            in.trap_box[k] = fi.trap_box[k];
        }
        in.trap_radius = fi.trap_radius;
        in.trap_tube = fi.trap_tube;
        hoist(in);
        // image slots: instance * 4 for its own formula, + 1 + j for mix-in j (template.wgsl:1341-1356)
        if (in.formula == ND_FORMULA_HEIGHTMAP) nd_image_hoist(c, 4 * i, in.h);
        // hybrid mix-ins (eval_mixed, template.wgsl:1361-1389; hybrid_cycle_len, :1332-1338)
        in.hybrid = fi.mixin_count > 0.5f;
        if (in.hybrid) {
            in.mix_count = std::min(std::max(i32(fi.mixin_count), 0), ND_MAX_MIXINS);
            // Cycles of two or more mix-ins are switched off: a GPU render of one (Julia bulb
            // base, Sierpinski and Mandelbulb mix-ins) did not finish, with no cause found in
            // the kernels or on a host run of the same per-ray code, which finished and matched.
            if (in.mix_count > ND_HIP_MAX_ACTIVE_MIXINS && i < s.count && !in.hidden)
                return nd_fail(c, ND_E_UNSUPPORTED_FORMULA, "instance %d: hybrid cycles of %d mix-ins are switched off (at most %d)",
                               i, in.mix_count, ND_HIP_MAX_ACTIVE_MIXINS);
            in.hyb_base_n = std::max(i32(fi.hybrid_base_iters), 0);
            in.hyb_total = std::max(i32(fi.hybrid_total_iters), 0);
            int cycle = in.hyb_base_n;
            DevMixin& base = in.mix[0];
            base.formula = in.formula;
            base.iters = in.hyb_base_n;
            for (int k = 0; k < 16; ++k) { base.p[k] = in.p[k]; base.dparam[k] = -1; }
            hoist(base);
            if (base.formula == ND_FORMULA_HEIGHTMAP) nd_image_hoist(c, 4 * i, base.h);
            for (int j = 0; j < in.mix_count; ++j) {
                const nd_mixin_params& mp = fi.mixins[j];
                DevMixin& m = in.mix[1 + j];
                const int mcode = i32(mp.code);
                m.formula = (mcode >= 0 && mcode < ND_MAX_CODES) ? c->code_map[mcode] : -1;
                if (i < s.count && !in.hidden && m.formula < 0)
                    return nd_fail(c, ND_E_UNSUPPORTED_FORMULA, "instance %d mix-in %d: formula code %d has no HIP distance estimator bound", i, j, mcode);
                m.iters = std::max(i32(mp.iterations), 0);
                cycle += m.iters;
                for (int k = 0; k < 16; ++k) { m.p[k] = mp.params[k]; m.dparam[k] = -1; }
                hoist(m);
                if (m.formula == ND_FORMULA_HEIGHTMAP) nd_image_hoist(c, 4 * i + 1 + j, m.h);
            }
            in.hyb_cycle = std::max(cycle, 1);
        }
    }
    return ND_OK;
}

nd_status nd_pack_job(nd_ctx* c, const nd_uniforms& u, const nd_march_opts& o, nd_rect rect, DevJob& j) {
    nd_status st = nd_pack_scene(c, u, j.scene);
    if (st != ND_OK) return st;
    if (!(u.resolution[0] >= 1.0f && u.resolution[1] >= 1.0f && u.resolution[0] <= 65536.0f && u.resolution[1] <= 65536.0f))
        return nd_fail(c, ND_E_INVALID, "resolution %gx%g out of range", u.resolution[0], u.resolution[1]);
    j.target_w = (int)u.resolution[0];
    j.target_h = (int)u.resolution[1];
    if (rect.w == 0 || rect.h == 0) {
        rect.x = rect.y = 0;
        rect.w = (uint32_t)j.target_w;
        rect.h = (uint32_t)j.target_h;
    }
    if ((uint64_t)rect.x + rect.w > (uint64_t)j.target_w || (uint64_t)rect.y + rect.h > (uint64_t)j.target_h)
        return nd_fail(c, ND_E_INVALID, "rect %ux%u+%u+%u outside the %dx%d target", rect.w, rect.h, rect.x, rect.y, j.target_w, j.target_h);
    j.rect_x = (int)rect.x; j.rect_y = (int)rect.y; j.rect_w = (int)rect.w; j.rect_h = (int)rect.h;
    // tiles of 2^tile_lg x 64 / 2^tile_lg pixels, one wave64 each (8 x 8 unless tuned)
    const uint32_t tw = c->tuning.tile_width ? c->tuning.tile_width : 8u;
    j.tile_lg = tw == 4 ? 2 : tw == 16 ? 4 : tw == 32 ? 5 : 3;
    const int th = 64 >> j.tile_lg;
    j.tiles_x = (j.rect_w + (1 << j.tile_lg) - 1) >> j.tile_lg;
    j.tiles_y = (j.rect_h + th - 1) / th;
    j.ray_count = (uint32_t)j.tiles_x * (uint32_t)j.tiles_y * 64u;
    for (int k = 0; k < 3; ++k) {
        j.cam[k] = u.camera_pos[k];
        j.right[k] = u.camera_right[k];
        j.up[k] = u.camera_up[k];
        j.fwd[k] = u.camera_forward[k];
    }
    j.tile_scale[0] = u.tile_scale[0]; j.tile_scale[1] = u.tile_scale[1];
    j.tile_bias[0] = u.tile_bias[0]; j.tile_bias[1] = u.tile_bias[1];
    j.aspect = u.resolution[0] / u.resolution[1];
    j.inv_tw = 1.0f / (float)j.target_w;
    j.inv_th = 1.0f / (float)j.target_h;
    // march_to preamble (template.wgsl:1859-1871), coarse and overstep modes off
    j.max_steps = std::max(i32(u.max_steps), 0);
    j.max_dist = u.max_dist;
    j.eps_c = u.epsilon_coefficient;
    j.eps_f = u.epsilon_floor;
    j.hq = u.high_quality > 0.5f;
    int budget = std::max(j.hq ? i32(u.refine_hq) : i32(u.refine_fast), 0);
    bool on_hit = u.march_verify < 0.5f;
    j.refine_budget = on_hit ? 0 : budget;
    j.bisect_budget = on_hit ? budget : 0;
    j.footprint = u.hq_footprint_budget_px * 2.0f / std::max(u.resolution[1], 1.0f);
    // light 0 soft shadow (shadow_factor, template.wgsl:2365-2375)
    j.shadow_on = 0;
    const nd_light& L = u.lights[0];
    if (o.shadows && i32(u.light_count) >= 1 && L.cast_shadows > 0.5f) {
        j.shadow_on = 1;
        j.light_type = i32(L.light_type + 0.5f) == 1 ? 1 : 0;
        if (j.light_type == 1) {
            double n = std::sqrt((double)L.position_or_direction[0] * L.position_or_direction[0] +
                                 (double)L.position_or_direction[1] * L.position_or_direction[1] +
                                 (double)L.position_or_direction[2] * L.position_or_direction[2]);
            if (!(n > 0)) return nd_fail(c, ND_E_INVALID, "light 0 has a zero direction");
            for (int k = 0; k < 3; ++k) j.light_v[k] = (float)(L.position_or_direction[k] / n);
        } else {
            for (int k = 0; k < 3; ++k) j.light_v[k] = L.position_or_direction[k];
        }
        j.light_k = L.shadow_softness + (512.0f - L.shadow_softness) * L.hard_shadows;
        j.light_far = u.max_dist * 2.0f;
        j.shadow_steps = o.shadow_max_steps ? (int)o.shadow_max_steps : std::min(i32(u.max_steps), 40);
    }
    j.normal_mode = (int)o.normal_mode;
    // empty-space accelerator: the grid is built by nd_accel_prepare
    j.accel = nullptr;
    j.accel_res = 0;
    j.accel_levels = 0;
    if (u.accel_enabled > 0.5f) {
        const int res = i32(u.accel_res), levels = std::min(i32(u.accel_levels), ND_MAX_CASCADES);
        if (res < 1 || res > 1024 || levels < 1)
            return nd_fail(c, ND_E_INVALID, "accelerator of %d^3 cells x %d levels", res, levels);
        j.accel_res = res;
        j.accel_levels = levels;
        j.accel_safety = u.accel_safety;
        for (int l = 0; l < ND_MAX_CASCADES; ++l)
            for (int k = 0; k < 4; ++k) j.accel_prm[l][k] = u.accel_params[l][k];
    }
    return ND_OK;
}

// ----------------------------------------------------------------------------- context
extern "C" {

nd_status nd_ctx_create(int hip_device, nd_ctx** out) {
    if (!out) return ND_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return ND_E_DEVICE;
    if (hip_device < 0 || hip_device >= n) return ND_E_INVALID;
    if (hipSetDevice(hip_device) != hipSuccess) return ND_E_DEVICE;
    nd_ctx* c = new nd_ctx();
    c->device = hip_device;
    c->noalloc_scratch = nd_noalloc_scratch(hip_device);  // device address of this GPU's scratch array
    for (int i = 0; i < ND_MAX_CODES; ++i) c->code_map[i] = i < 4 ? i : -1;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return ND_E_DEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, hip_device) == hipSuccess) {
        c->cu_count = prop.multiProcessorCount;
        c->mem_total = prop.totalGlobalMem;
        std::snprintf(c->arch, sizeof(c->arch), "%s", prop.gcnArchName);
    }
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, hip_device) == hipSuccess && khz > 0)
        c->wall_khz = khz;
    for (auto& e : c->ev) (void)hipEventCreate(&e);
    nd_resident_march(c->resident);
    nd_resident_adjoint(c->resident);
    nd_resident_mc(c->resident);
    c->tuning.waves_per_cu = 0;  // 0 = each kernel's resident waves per CU, fewer for small batches
    c->tuning.refill_min_idle = 8;
    c->tuning.chunk_rays = 0;  // 0 = 96 rays per claim with a shared scene, 256 without
    c->tuning.flags = 0;
    *out = c;
    return ND_OK;
}

void nd_ctx_destroy(nd_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    nd_drop_graphs(c);
    for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
    c->jobs.release(); c->job_base.release(); c->ctr.release(); c->hits.release(); c->shared.release();
    c->staging.release(); c->latency.release(); c->adj.release(); c->mc.release();
    c->accel.release(); c->accel_scenes.release(); c->mc_adapt.release(); c->color_scenes.release();
    c->image_buf.release();
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* nd_last_error(const nd_ctx* c) { return c ? c->err : "null context"; }

nd_status nd_sync(nd_ctx* c) {
    if (!c) return ND_E_INVALID;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return nd_hip_fail(c, e, "hipStreamSynchronize");
    return ND_OK;
}

nd_status nd_set_cancel(nd_ctx* c, const volatile int* flag) {
    if (!c) return ND_E_INVALID;
    c->cancel = flag;
    return ND_OK;
}

void* nd_ctx_stream(nd_ctx* c) { return c ? (void*)c->stream : nullptr; }

nd_status nd_bind_formulas(nd_ctx* c, const int32_t* map, uint32_t n) {
    if (!c) return ND_E_INVALID;
    if ((!map && n) || n > ND_MAX_CODES) return nd_fail(c, ND_E_INVALID, "code table of %u entries (max %d)", n, ND_MAX_CODES);
    for (uint32_t i = 0; i < n; ++i)
        if (map[i] < -1 || map[i] >= ND_FORMULA_COUNT)
            return nd_fail(c, ND_E_INVALID, "code %u maps to unknown built-in %d", i, map[i]);
    for (int i = 0; i < ND_MAX_CODES; ++i) c->code_map[i] = (uint32_t)i < n ? map[i] : -1;
    return ND_OK;
}

nd_status nd_get_tuning(nd_ctx* c, nd_tuning* out) {
    if (!c || !out) return ND_E_INVALID;
    *out = c->tuning;
    return ND_OK;
}

nd_status nd_set_tuning(nd_ctx* c, const nd_tuning* t) {
    if (!c || !t) return ND_E_INVALID;
    if (t->refill_min_idle < 1 || t->refill_min_idle > 64) return nd_fail(c, ND_E_INVALID, "refill_min_idle must be 1..64");
    if (t->chunk_rays > (1u << 14)) return nd_fail(c, ND_E_INVALID, "chunk_rays must be 0..2^14");
    if (t->waves_per_cu > 64) return nd_fail(c, ND_E_INVALID, "waves_per_cu must be 0..64");
    if (t->flags & ~(ND_TUNE_TRACK_LATENCY | ND_TUNE_NO_GRAPH | ND_TUNE_GBUFFER_GRAPH)) return nd_fail(c, ND_E_INVALID, "unknown tuning flags 0x%x", t->flags);
    if (t->tile_width != 0 && t->tile_width != 4 && t->tile_width != 8 && t->tile_width != 16 && t->tile_width != 32)
        return nd_fail(c, ND_E_INVALID, "tile_width must be 0, 4, 8, 16 or 32");
    if (t->refill_period > 16) return nd_fail(c, ND_E_INVALID, "refill_period must be 0..16");
    if (t->mc_pool_log2 != 0 && (t->mc_pool_log2 < 16 || t->mc_pool_log2 > 27))
        return nd_fail(c, ND_E_INVALID, "mc_pool_log2 must be 0 or 16..27");
    c->tuning = *t;
    c->round_cap = t->round_cap ? t->round_cap : 0xffffffffu;
    return ND_OK;
}

}  // extern "C"

// ----------------------------------------------------------------------------- batch
static size_t out_bytes(const nd_gbuffer_out& o, size_t px) {
    size_t b = 0;
    if (o.depth) b += 4 * px;
    if (o.normal) b += 12 * px;
    if (o.position) b += 12 * px;
    if (o.trap) b += 4 * px;
    if (o.hit) b += px;
    if (o.shadow) b += 4 * px;
    if (o.evals) b += 4 * px;
    if (o.color) b += 12 * px;
    if (o.albedo) b += 12 * px;
    // the colour pass reads the hit and depth of every pixel: internal buffers when not asked for
    if ((o.color || o.albedo) && !o.depth) b += 4 * px;
    if ((o.color || o.albedo) && !o.hit) b += px;
    return b;
}

uint32_t nd_pass_grid(const nd_ctx* c, int res, uint64_t units) {
    const uint64_t cu = (uint64_t)std::max(c->cu_count, 1);
    uint64_t wpc = c->tuning.waves_per_cu;
    if (wpc == 0) {
        // As many waves as fit at once: a wave beyond that waits for a slot and then finds
        // the queue drained, and more waves per SIMD only slow the resident ones. A small
        // batch gets about two claims per wave (at least 4 waves per CU), which measured
        // faster than one.
        wpc = c->resident[res] ? c->resident[res] : 8;
        if (units) wpc = std::min<uint64_t>(wpc, std::max<uint64_t>(4, (units + 2 * cu - 1) / (2 * cu)));
    }
    return (uint32_t)(cu * wpc);
}

// Tangent slots of the adjoint: 0..2 spatial, 3 Mandelbulb power, 4 Mandelbox scale,
// 5 Mandelbox min radius, on the first visible instance of each formula.
static void attach_adjoint_slots(DevScene& s) {
    bool bulb = false, box = false;
    for (int i = 0; i < s.count; ++i) {
        DevInst& in = s.inst[i];
        if (in.hidden) continue;
        if (in.formula == ND_FORMULA_MANDELBULB && !bulb) { in.dparam[0] = 3; bulb = true; }
        if (in.formula == ND_FORMULA_MANDELBOX && !box) { in.dparam[0] = 4; in.dparam[1] = 5; box = true; }
        if (in.hybrid)  // a hybrid's base steps read their parameters from mix[0]
            for (int k = 0; k < 16; ++k) in.mix[0].dparam[k] = in.dparam[k];
    }
}

nd_status nd_batch_impl(nd_ctx* c, const nd_job* jobs, uint32_t n, const nd_march_opts* opts, nd_batch_stats* stats,
                        const NdBatchExtra* extra) {
    if (!c) return ND_E_INVALID;
    NdBatchExtra ex;
    if (extra) ex = *extra;
    const NdFixedWorkspace* fx = ex.fixed;  // non-null: carve every buffer from it, allocate nothing
    if (!jobs || n == 0) return nd_fail(c, ND_E_INVALID, "empty batch");
    if (n > 65535) return nd_fail(c, ND_E_INVALID, "at most 65535 jobs per batch");
    if (fx && n != 1) return nd_fail(c, ND_E_INVALID, "the fixed workspace holds one job");
    if (c->cancel && *c->cancel) return nd_fail(c, ND_E_CANCELLED, "cancelled before launch");
    nd_march_opts o{};
    if (opts) o = *opts;
    if (o.normal_mode > ND_NORMAL_CENTRAL) return nd_fail(c, ND_E_INVALID, "normal_mode %u", o.normal_mode);
    if (o.kernel > ND_KERNEL_PERSISTENT) return nd_fail(c, ND_E_INVALID, "kernel %u", o.kernel);
    if (hipSetDevice(c->device) != hipSuccess) return nd_fail(c, ND_E_DEVICE, "hipSetDevice");
    // launches below are checked with hipGetLastError: start from a clean slate, so an error
    // left by an earlier call that was not checked is not reported as this batch's
    (void)hipGetLastError();

    // host records: the context's vectors, or (fixed workspace, one job) the stack
    DevJob fx_job;
    uint32_t fx_base = 0;
    std::vector<uint32_t> base_vec;
    DevJob* hj;
    uint32_t* base;
    if (fx) {
        std::memset(&fx_job, 0, sizeof(fx_job));
        hj = &fx_job;
        base = &fx_base;
    } else {
        c->host_jobs.assign(n, DevJob{});
        base_vec.assign(n, 0u);
        hj = c->host_jobs.data();
        base = base_vec.data();
    }
    uint64_t total = 0;
    size_t staging = 0;
    bool any_shadow = false, any_accel = false, any_color = false;
    for (uint32_t i = 0; i < n; ++i) {
        if (!jobs[i].uniforms) return nd_fail(c, ND_E_INVALID, "job %u: null uniforms", i);
        nd_status st = nd_pack_job(c, *jobs[i].uniforms, o, jobs[i].rect, hj[i]);
        if (st != ND_OK) return st;
        if (ex.adjoint_slots) attach_adjoint_slots(hj[i].scene);
        hj[i].ray_base = (uint32_t)total;
        base[i] = (uint32_t)total;
        total += hj[i].ray_count;
        any_shadow = any_shadow || hj[i].shadow_on;
        any_accel = any_accel || hj[i].accel_levels > 0;
        const nd_gbuffer_out& go = jobs[i].out;
        size_t px = (size_t)hj[i].rect_w * hj[i].rect_h;
        const bool color = go.color || go.albedo;
        any_color = any_color || color;
        if (color) {
            const uint32_t parts = (uint32_t)jobs[i].uniforms->debug_parts;
            if (parts & ((1u << 17) | (1u << 18)))  // PART_HOTSPOTS, PART_OVERSTEP
                return nd_fail(c, ND_E_UNSUPPORTED, "job %u: the hotspot and overstep views are not implemented", i);
        }
        if (!go.device_ptrs) staging += out_bytes(go, px) + 9 * 256;
        else if (color && (!go.depth || !go.hit)) {
            if (fx) return nd_fail(c, ND_E_INVALID, "job %u: without allocation the colour outputs need the depth and hit outputs", i);
            staging += ((!go.depth ? 4 * px : 0) + (!go.hit ? px : 0)) + 2 * 256;
        }
    }
    if (total >= (1ull << 31)) return nd_fail(c, ND_E_INVALID, "batch of %llu ray slots exceeds 2^31", (unsigned long long)total);
    if (fx) {
        if (staging) return nd_fail(c, ND_E_INVALID, "without allocation the outputs must be device pointers (device_ptrs = 1)");
        if (any_accel)
            return nd_fail(c, ND_E_UNSUPPORTED, "without allocation the empty-space accelerator is not available (its grids are allocated)");
        if (total > fx->max_rays)
            return nd_fail(c, ND_E_INVALID, "%llu ray slots (64 per 8x8 tile of the rect) exceed the %u of the fixed workspace",
                           (unsigned long long)total, fx->max_rays);
    }

    // outputs: device pointers as given, host pointers through a staging buffer
    nd_status st = ND_OK;
    if (staging && (st = c->staging.ensure(c, staging)) != ND_OK) return st;
    {
        char* sp = (char*)c->staging.p;
        auto carve = [&](size_t bytes) { char* r = sp; sp += (bytes + 255) & ~(size_t)255; return r; };
        for (uint32_t i = 0; i < n; ++i) {
            const nd_gbuffer_out& go = jobs[i].out;
            DevJob& j = hj[i];
            size_t px = (size_t)j.rect_w * j.rect_h;
            const bool color = go.color || go.albedo;
            if (go.device_ptrs) {
                j.out_depth = go.depth; j.out_normal = go.normal; j.out_pos = go.position; j.out_trap = go.trap;
                j.out_hit = go.hit; j.out_shadow = go.shadow; j.out_evals = go.evals;
                j.out_color = go.color; j.out_albedo = go.albedo;
            } else {
                j.out_depth = go.depth ? (float*)carve(4 * px) : nullptr;
                j.out_normal = go.normal ? (float*)carve(12 * px) : nullptr;
                j.out_pos = go.position ? (float*)carve(12 * px) : nullptr;
                j.out_trap = go.trap ? (float*)carve(4 * px) : nullptr;
                j.out_hit = go.hit ? (uint8_t*)carve(px) : nullptr;
                j.out_shadow = go.shadow ? (float*)carve(4 * px) : nullptr;
                j.out_evals = go.evals ? (uint32_t*)carve(4 * px) : nullptr;
                j.out_color = go.color ? (float*)carve(12 * px) : nullptr;
                j.out_albedo = go.albedo ? (float*)carve(12 * px) : nullptr;
            }
            if (color && !j.out_depth) j.out_depth = (float*)carve(4 * px);
            if (color && !j.out_hit) j.out_hit = (uint8_t*)carve(px);
        }
    }

    bool shared = true;
    for (uint32_t i = 1; i < n && shared; ++i) shared = std::memcmp(&hj[i].scene, &hj[0].scene, sizeof(DevScene)) == 0;
    bool ext = false;
    for (uint32_t i = 0; i < n && !ext; ++i) ext = nd_scene_ext(hj[i].scene);
    c->last_shared = shared && !ext;  // the EXT kernels read each job's own scene
    c->last_ext = ext;
    if (!fx && (st = nd_accel_prepare(c, hj, n)) != ND_OK) return st;

    // device records: grown context buffers, or carved from the fixed workspace
    const bool track = !fx && (c->tuning.flags & ND_TUNE_TRACK_LATENCY) != 0;
    const size_t lat_jobs = (sizeof(unsigned long long) * n + 255) & ~(size_t)255;
    char* fxp = fx ? fx->base : nullptr;
    char* const fx_end = fx ? fx->base + fx->bytes : nullptr;
    auto fcarve = [&](size_t bytes) -> char* {
        char* r = fxp;
        fxp += (bytes + 255) & ~(size_t)255;
        return fxp <= fx_end ? r : nullptr;
    };
    DevJob* d_jobs;
    uint32_t* d_job_base;
    DevCounters* d_ctr;
    DevScene* d_shared;
    char* d_latency = nullptr;
    if (fx) {
        d_jobs = (DevJob*)fcarve(sizeof(DevJob) * n);
        d_job_base = (uint32_t*)fcarve(sizeof(uint32_t) * n);
        d_ctr = (DevCounters*)fcarve(sizeof(DevCounters));
        d_shared = (DevScene*)fcarve(sizeof(DevScene));
        if (!d_jobs || !d_job_base || !d_ctr || !d_shared) return nd_fail(c, ND_E_INVALID, "fixed workspace too small for the job records");
    } else {
        if ((st = c->jobs.ensure(c, sizeof(DevJob) * n)) != ND_OK) return st;
        if ((st = c->job_base.ensure(c, sizeof(uint32_t) * n)) != ND_OK) return st;
        if ((st = c->ctr.ensure(c, sizeof(DevCounters))) != ND_OK) return st;
        if ((st = c->shared.ensure(c, sizeof(DevScene))) != ND_OK) return st;
        // latency: per-job completion clocks, then (when tracking) one completion tick per ray slot
        if ((st = c->latency.ensure(c, lat_jobs + (track ? sizeof(uint32_t) * (size_t)total : 0))) != ND_OK) return st;
        d_jobs = (DevJob*)c->jobs.p;
        d_job_base = (uint32_t*)c->job_base.p;
        d_ctr = (DevCounters*)c->ctr.p;
        d_shared = (DevScene*)c->shared.p;
        d_latency = (char*)c->latency.p;
    }
    // per-job material scenes of the colour pass
    char* d_color_scenes = nullptr;
    std::vector<unsigned char> color_scenes;
    alignas(16) unsigned char fx_color_scene[ND_MC_SCENE_MAX_BYTES];
    const size_t mcs = nd_mc_scene_bytes();
    if (any_color) {
        unsigned char* hs;
        if (fx) {
            if (!(d_color_scenes = fcarve(ND_MC_SCENE_MAX_BYTES))) return nd_fail(c, ND_E_INVALID, "fixed workspace too small for the colour pass");
            hs = fx_color_scene;
        } else {
            if ((st = c->color_scenes.ensure(c, mcs * n)) != ND_OK) return st;
            d_color_scenes = (char*)c->color_scenes.p;
            color_scenes.resize(mcs * n);
            hs = color_scenes.data();
        }
        for (uint32_t i = 0; i < n; ++i) nd_mc_scene_pack(*jobs[i].uniforms, hs + mcs * i);
    }
    hipError_t e;
    if ((e = hipMemcpyAsync(d_jobs, hj, sizeof(DevJob) * n, hipMemcpyHostToDevice, c->stream)) != hipSuccess)
        return nd_hip_fail(c, e, "upload jobs");
    if (any_color &&
        (e = hipMemcpyAsync(d_color_scenes, fx ? fx_color_scene : color_scenes.data(), mcs * n, hipMemcpyHostToDevice,
                            c->stream)) != hipSuccess)
        return nd_hip_fail(c, e, "upload material scenes");
    if ((e = hipMemcpyAsync(d_job_base, base, sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream)) != hipSuccess)
        return nd_hip_fail(c, e, "upload job table");
    if (shared && (e = hipMemcpyAsync(d_shared, &hj[0].scene, sizeof(DevScene), hipMemcpyHostToDevice, c->stream)) != hipSuccess)
        return nd_hip_fail(c, e, "upload scene");

    // a failure after the uploads were queued waits for them: their source may be this frame
    auto drained = [&](nd_status r) { (void)hipStreamSynchronize(c->stream); return r; };
    nd_batch_stats s{};
    float ms_total = 0, ms_march = 0, ms_surface = 0, ms_shadow = 0;
    const bool persistent = o.kernel != ND_KERNEL_PLAIN;
    if (!persistent) {
        // one launch per frame: the job record is read from device memory by one kernel each
        if ((e = nd_launch_zero_counters(d_ctr, c->stream)) != hipSuccess)
            return drained(nd_hip_fail(c, e, "zero counters"));
        (void)hipEventRecord(c->ev[0], c->stream);
        for (uint32_t i = 0; i < n; ++i) {
            e = nd_launch_plain(d_jobs + i, hj[i].tiles_x, hj[i].tiles_y, d_ctr, ext, c->stream);
            if (e != hipSuccess) return drained(nd_hip_fail(c, e, "launch plain march"));
        }
        (void)hipEventRecord(c->ev[3], c->stream);
        s.kernels = n;
    } else {
        // hit slots: every ray may hit, plus at most one partly used 64-slot chunk per wave
        const bool acc = any_accel;
        // Claim size: with one scene for every job (the 16-camera batch), 96-ray claims give the
        // march pass the finest tail (-4.5% against 256); a batch of different scenes changes
        // job and scene at more claims and is fastest at 256 (the 256-job batch: 62 ms at 256,
        // 65-69 ms at 96-128).
        const uint32_t chunk = c->tuning.chunk_rays ? c->tuning.chunk_rays : (shared && !ext ? 96u : 256u);
        const uint64_t claims = ((uint64_t)total + chunk - 1) / chunk;
        const int res_march = ext ? (acc ? RES_MARCH_ACC_EXT : RES_MARCH_EXT)
                                  : acc ? (shared ? RES_MARCH_ACC : RES_MARCH_ACC_MIXED) : (shared ? RES_MARCH : RES_MARCH_MIXED);
        const int res_shadow = ext ? (acc ? RES_SHADOW_ACC_EXT : RES_SHADOW_EXT)
                                   : acc ? (shared ? RES_SHADOW_ACC : RES_SHADOW_ACC_MIXED) : (shared ? RES_SHADOW : RES_SHADOW_MIXED);
        const int res_surface = ext ? RES_SURFACE_EXT : (shared ? RES_SURFACE : RES_SURFACE_MIXED);
        const uint32_t grid_cap = fx ? fx->max_grid : 0xffffffffu;
        const uint32_t grid_march = std::min(nd_pass_grid(c, res_march, claims), grid_cap);
        const uint32_t grid_shadow = std::min(nd_pass_grid(c, res_shadow, claims), grid_cap);
        const uint32_t grid_surface = std::min(nd_pass_grid(c, res_surface, ((uint64_t)total + 63) / 64), grid_cap);
        // the claim counters run past the end by at most one claim per wave (shadow claims
        // are 4 x chunk_rays); they must not wrap
        const uint32_t grid_max = std::max(grid_march, grid_shadow);
        if ((uint64_t)total + (uint64_t)grid_max * 4u * chunk >= (1ull << 32))
            return drained(nd_fail(c, ND_E_INVALID, "batch of %llu rays too large for one persistent launch",
                                   (unsigned long long)total));
        size_t hit_cap = (size_t)total + (size_t)grid_march * 64;
        size_t hb_bytes = hit_cap * (4 + 2 + 4 + 4 + 2 + 7 * 4) + 16 * 256;
        char* hp;
        if (fx) {
            if (!(hp = fcarve(hb_bytes)))
                return drained(nd_fail(c, ND_E_INVALID, "fixed workspace too small for %zu hit slots", hit_cap));
        } else {
            if ((st = c->hits.ensure(c, hb_bytes)) != ND_OK) return drained(st);
            hp = (char*)c->hits.p;
        }
        auto carve = [&](size_t bytes) { char* r = hp; hp += (bytes + 255) & ~(size_t)255; return r; };
        NdPassLaunch L;
        std::memset(&L, 0, sizeof(L));  // compared bytewise as the graph key: no stray padding
        L.hb.ray = (uint32_t*)carve(4 * hit_cap);
        L.hb.job = (uint16_t*)carve(2 * hit_cap);
        L.hb.t = (float*)carve(4 * hit_cap);
        L.hb.opix = (uint32_t*)carve(4 * hit_cap);
        L.hb.steps = (uint16_t*)carve(2 * hit_cap);
        L.hb.sox = (float*)carve(4 * hit_cap);
        L.hb.soy = (float*)carve(4 * hit_cap);
        L.hb.soz = (float*)carve(4 * hit_cap);
        L.hb.sdx = (float*)carve(4 * hit_cap);
        L.hb.sdy = (float*)carve(4 * hit_cap);
        L.hb.sdz = (float*)carve(4 * hit_cap);
        L.hb.sdist = (float*)carve(4 * hit_cap);
        L.hb.capacity = (uint32_t)hit_cap;
        L.jobs = d_jobs;
        L.shared = shared ? d_shared : nullptr;
        L.job_base = d_job_base;
        L.n_jobs = (int)n;
        L.total_rays = (uint32_t)total;
        L.ctr = d_ctr;
        L.grid_march = grid_march;
        L.grid_surface = grid_surface;
        L.grid_shadow = grid_shadow;
        c->last_rays = (uint32_t)total;
        c->last_hb = L.hb;
        L.chunk = chunk;
        L.min_idle = c->tuning.refill_min_idle;
        L.round_cap = c->round_cap;
        L.refill_period = c->tuning.refill_period ? c->tuning.refill_period : 1u;
        L.any_shadow = any_shadow;
        L.surface = !ex.skip_surface;
        L.accel = acc;
        L.ext = ext;
        L.job_done_cycles = track ? (unsigned long long*)d_latency : nullptr;
        L.ray_done = track ? (uint32_t*)(d_latency + lat_jobs) : nullptr;
        // The passes of a batch shape are captured into a hipGraph when the same launch record -
        // grids, claims, flags and every device pointer - comes a second time, and replayed
        // while it stays the same. The first batch of a shape runs pass by pass, which also
        // loads the kernels' code before any capture. The job records are uploaded before each
        // replay. No memset node: the counters are cleared by a kernel. The timing events stay
        // out of the graph (an event recorded only by a graph node has no time to read back):
        // a replay is timed as a whole, between events recorded around the graph launch.
        const bool graphs = !fx && !c->graph_broken && (c->tuning.flags & ND_TUNE_GBUFFER_GRAPH);
        bool replayed = false;
        if (graphs) {
            const bool same = c->graph_exec && std::memcmp(&L, &c->graph_key, sizeof(L)) == 0;
            const bool repeat = std::memcmp(&L, &c->graph_seen, sizeof(L)) == 0;
            if (!same && !repeat) std::memcpy(&c->graph_seen, &L, sizeof(L));
            if (!same && repeat) {
                if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
                c->graph_exec = nullptr;
                hipGraph_t g = nullptr;
                e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
                if (e == hipSuccess) {
                    hipError_t el = nd_launch_persistent(L, c->stream);
                    e = hipStreamEndCapture(c->stream, &g);
                    if (e == hipSuccess) e = el;
                }
                if (e == hipSuccess) e = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
                if (g) (void)hipGraphDestroy(g);
                if (e == hipSuccess) {
                    std::memcpy(&c->graph_key, &L, sizeof(L));
                } else {
                    // capture unavailable: leave the stream out of capture mode and launch pass by pass
                    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
                    if (hipStreamIsCapturing(c->stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
                        hipGraph_t g2 = nullptr;
                        (void)hipStreamEndCapture(c->stream, &g2);
                        if (g2) (void)hipGraphDestroy(g2);
                    }
                    (void)hipGetLastError();
                    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
                    c->graph_exec = nullptr;
                    c->graph_broken = true;
                }
            }
            if (c->graph_exec) {
                (void)hipEventRecord(c->ev[0], c->stream);
                e = hipGraphLaunch(c->graph_exec, c->stream);
                if (e != hipSuccess) return drained(nd_hip_fail(c, e, "replay persistent passes"));
                (void)hipEventRecord(c->ev[3], c->stream);
                replayed = true;
            }
        }
        if (!replayed) {
            for (int k = 0; k < 4; ++k) L.ev[k] = c->ev[k];
            e = nd_launch_persistent(L, c->stream);
            if (e != hipSuccess) return drained(nd_hip_fail(c, e, "launch persistent passes"));
        }
        s.graph = replayed ? 1u : 0u;
        s.kernels = 3 + (any_shadow ? 1 : 0) + 1;
    }
    if (any_color) {
        // fs_simple colour and fs_aov albedo from the hit and depth just written
        if ((e = nd_launch_gbuffer_color(d_jobs, d_color_scenes, d_job_base, (int)n, (uint32_t)total, ext, c->stream)) != hipSuccess)
            return drained(nd_hip_fail(c, e, "launch colour pass"));
        (void)hipEventRecord(c->ev[4], c->stream);
        s.kernels += 1;
    }
    e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return nd_hip_fail(c, e, "batch execution");
    // a failed time read leaves no error behind for the next call's launch checks
    auto elapsed = [&](float* ms, hipEvent_t a, hipEvent_t b) {
        if (hipEventElapsedTime(ms, a, b) != hipSuccess) { *ms = 0.0f; (void)hipGetLastError(); }
    };
    elapsed(&ms_total, c->ev[0], c->ev[any_color ? 4 : 3]);
    DevCounters dc;
    if ((e = hipMemcpy(&dc, d_ctr, sizeof(dc), hipMemcpyDeviceToHost)) != hipSuccess)
        return nd_hip_fail(c, e, "read counters");
    if (persistent) {
        if (s.graph) {
            elapsed(&ms_march, c->ev[0], c->ev[3]);  // the replayed passes as one
        } else {
            elapsed(&ms_march, c->ev[0], c->ev[1]);
            elapsed(&ms_surface, c->ev[1], c->ev[2]);
            elapsed(&ms_shadow, c->ev[2], c->ev[3]);
        }
        s.hits = dc.hits;
        s.de_evals_march = dc.de_march;
        s.march_wave_rounds = dc.march_rounds;
        s.de_evals_surface = dc.de_surface;
        s.de_evals_shadow = dc.de_shadow;
        s.de_evals = dc.de_march + dc.de_surface + dc.de_shadow;
        if (!fx) {
            c->job_latency_ms.assign(n, 0.0f);
            if (track) {
                std::vector<unsigned long long> done(n);
                if ((e = hipMemcpy(done.data(), d_latency, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost)) != hipSuccess)
                    return nd_hip_fail(c, e, "read latency");
                unsigned long long t0 = dc.clock_start;
                for (uint32_t i = 0; i < n; ++i)
                    c->job_latency_ms[i] = done[i] > t0 ? (float)((double)(done[i] - t0) / (double)c->wall_khz) : 0.0f;
            }
        }
        if (dc.rounds_capped) return nd_fail(c, ND_E_TIMEOUT, "%u waves reached the round cap", dc.rounds_capped);
    } else {
        s.de_evals = dc.de_march;
        s.de_evals_march = dc.de_march;
        if (!fx) c->job_latency_ms.assign(n, ms_total);
    }
    for (uint32_t i = 0; i < n; ++i) s.rays += (uint64_t)hj[i].rect_w * hj[i].rect_h;
    s.ms_total = ms_total;
    s.ms_march = ms_march;
    s.ms_surface = ms_surface;
    s.ms_shadow = ms_shadow;

    // host outputs back
    for (uint32_t i = 0; i < n; ++i) {
        const nd_gbuffer_out& go = jobs[i].out;
        if (go.device_ptrs) continue;
        const DevJob& j = hj[i];
        size_t px = (size_t)j.rect_w * j.rect_h;
        struct { void* dst; const void* src; size_t b; } cp[9] = {
            {go.depth, j.out_depth, 4 * px}, {go.normal, j.out_normal, 12 * px}, {go.position, j.out_pos, 12 * px},
            {go.trap, j.out_trap, 4 * px},
            {go.hit, j.out_hit, px}, {go.shadow, j.out_shadow, 4 * px}, {go.evals, j.out_evals, 4 * px},
            {go.color, j.out_color, 12 * px}, {go.albedo, j.out_albedo, 12 * px}};
        for (auto& k : cp) {
            if (!k.dst) continue;
            if ((e = hipMemcpy(k.dst, k.src, k.b, hipMemcpyDeviceToHost)) != hipSuccess) return nd_hip_fail(c, e, "read outputs");
        }
    }
    if (stats) *stats = s;
    if (c->cancel && *c->cancel) return nd_fail(c, ND_E_CANCELLED, "cancelled");
    return ND_OK;
}

extern "C" {

nd_status nd_render_batch(nd_ctx* c, const nd_job* jobs, uint32_t n, const nd_march_opts* opts, nd_batch_stats* stats) {
    return nd_batch_impl(c, jobs, n, opts, stats);
}

nd_status nd_render_gbuffer(nd_ctx* c, const void* uniforms, size_t size, const nd_march_opts* opts, nd_rect rect,
                            const nd_gbuffer_out* out, nd_batch_stats* stats) {
    if (!c) return ND_E_INVALID;
    if (!uniforms) return nd_fail(c, ND_E_INVALID, "null uniforms");
    if (size != sizeof(nd_uniforms)) return nd_fail(c, ND_E_INVALID, "uniforms_size %zu != %zu", size, sizeof(nd_uniforms));
    if (!out) return nd_fail(c, ND_E_INVALID, "null output set");
    nd_job j;
    j.uniforms = (const nd_uniforms*)uniforms;
    j.rect = rect;
    j.out = *out;
    return nd_batch_impl(c, &j, 1, opts, stats);
}

nd_status nd_batch_job_latency(nd_ctx* c, float* ms_out, uint32_t n) {
    if (!c || !ms_out) return ND_E_INVALID;
    if (n != c->job_latency_ms.size()) return nd_fail(c, ND_E_INVALID, "last batch had %zu jobs", c->job_latency_ms.size());
    std::memcpy(ms_out, c->job_latency_ms.data(), sizeof(float) * n);
    return ND_OK;
}

nd_status nd_autotune(nd_ctx* c, const nd_job* jobs, uint32_t n, const nd_march_opts* opts, nd_tuning* best) {
    if (!c) return ND_E_INVALID;
    nd_march_opts o{};
    if (opts) o = *opts;
    o.kernel = ND_KERNEL_PERSISTENT;
    const nd_tuning saved = c->tuning;
    const uint32_t occ = c->resident[RES_MARCH] ? c->resident[RES_MARCH] : 8;
    // candidates per parameter; 0 is each parameter's default rule
    const std::vector<uint32_t> cand[5] = {{0, std::max(occ / 2, 1u), occ, occ * 2},  // waves per CU
                                           {1, 8, 32},                                  // refill threshold
                                           {0, 96, 256, 1024},                          // claim size
                                           {8, 16, 32, 4},                              // tile width
                                           {1, 2, 4}};                                  // refill period (step budget)
    auto field = [](nd_tuning& t, int k) -> uint32_t& {
        switch (k) {
            case 0: return t.waves_per_cu;
            case 1: return t.refill_min_idle;
            case 2: return t.chunk_rays;
            case 3: return t.tile_width;
            default: return t.refill_period;
        }
    };
    auto time_of = [&](const nd_tuning& t, float* ms) -> nd_status {
        c->tuning = t;
        *ms = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {  // the best of two runs
            nd_batch_stats s;
            nd_status st = nd_batch_impl(c, jobs, n, &o, &s);
            if (st != ND_OK) return st;
            *ms = std::min(*ms, s.ms_total);
        }
        return ND_OK;
    };
    nd_tuning bt = saved;
    float bms;
    nd_status st = time_of(bt, &bms);
    if (st != ND_OK) { c->tuning = saved; return st; }
    for (int sweep = 0; sweep < 2; ++sweep)
        for (int k = 0; k < 5; ++k)
            for (uint32_t v : cand[k]) {
                nd_tuning t = bt;
                if (field(t, k) == v) continue;
                field(t, k) = v;
                float ms;
                if ((st = time_of(t, &ms)) != ND_OK) { c->tuning = saved; return st; }
                if (ms < bms) { bms = ms; bt = t; }
            }
    c->tuning = bt;
    if (best) *best = bt;
    return ND_OK;
}

}  // extern "C"
