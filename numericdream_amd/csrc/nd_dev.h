// nd_dev.h - device-side records shared by the host packer and the HIP kernels.
//
// The host decodes each 12,016-byte Uniforms block once (nd_pack.cpp) into a DevJob:
// the scene (instances with their hoisted per-formula constants), the camera, the march
// precision, light 0 and the output pointers. Kernels never parse Uniforms themselves.
#pragma once

#include <stdint.h>

#include "nd_uniforms.h"

#define ND_DEV_MAX_INST 4
#define ND_DEV_MAX_CARVES 16
// Mix-ins a hybrid cycle may use in the HIP kernels; cycles of more are rejected at packing
// (the Uniforms hold ND_MAX_MIXINS).
#define ND_HIP_MAX_ACTIVE_MIXINS 1

enum { NDF_BULB = 0, NDF_BOX = 1, NDF_MENGER = 2, NDF_KLEINIAN = 3 };

// A hybrid mix-in slot (MixinParams, fr.zig:112-117) after decoding: the formula steps of
// the slot read the same fields as an instance's own formula (p, h, rot_on, rot_cs).
struct DevMixin {
    int formula;       // nd_builtin_formula of the mix-in's code
    int iters;         // max(i32(iterations), 0): its steps per hybrid cycle
    int rot_on;
    int bulb_pow8;
    int dparam[16];    // -1: mix-in parameters carry no adjoint tangent
    float p[16];
    float h[16];
    float rot_cs[6];
    float _pad[2];
};

// One fractal instance (FractalInstance, fr.zig:119-145) after decoding.
struct DevInst {
    int formula;       // nd_builtin_formula, -1 = none (DE = 1e6)
    int combine;       // combine mode 0..6
    int hidden;
    int iters;         // i32(de_iterations(p))
    int identity;      // offset 0, scale 1, frame identity: the instance transform is skipped
    int rot_on;        // formula-internal Euler rotation (box / menger) is non-zero
    int bulb_pow8;     // Mandelbulb with power exactly 8: trig-free z^8
    int trap_shape;
    int trap_mode;
    int dparam[16];    // adjoint tangent slot of each @param (-1: constant)
    float blend_k;
    float de_factor;   // min|scale| * clamp(step_safety, .001, 1)
    float offset[3];
    float inv_s[3];    // 1 / signed scale magnitude
    float frame[9];    // rotation columns (c0.x c0.y c0.z c1.x ...)
    float p[16];       // raw @param values
    float h[16];       // hoisted per-formula constants (see nd_pack.cpp)
    float rot_cs[6];   // cos x, sin x, cos y, sin y, cos z, sin z of the formula rotation
    float trap_center[3];
    float trap_box[3];
    float trap_radius;
    float trap_tube;
    // hybrid cycle (eval_mixed, template.wgsl:1361-1389)
    int hybrid;        // mixin_count > 0.5: the instance runs eval_mixed
    int mix_count;     // min(i32(mixin_count), ND_MAX_MIXINS)
    int hyb_base_n;    // max(i32(hybrid_base_iters), 0)
    int hyb_total;     // max(i32(hybrid_total_iters), 0)
    int hyb_cycle;     // max(base_n + sum of the mix-ins' iters, 1)
    DevMixin mix[ND_MAX_MIXINS + 1];  // [0] the base formula's own slots, [1 + j] mix-in j
};

struct DevScene {
    int count;         // max(i32(instance_count), 1)
    int carve_count;
    float empty_de;    // max_dist * 2 + 1
    float _pad;
    DevInst inst[ND_DEV_MAX_INST];
    float carves[ND_DEV_MAX_CARVES][4];
};

// One render job: a frame (or rect of it) with its camera, precision, light and outputs.
struct DevJob {
    DevScene scene;
    float cam[3], right[3], up[3], fwd[3];
    float tile_scale[2], tile_bias[2];
    float aspect;
    float inv_tw, inv_th;  // 1 / target size
    int target_w, target_h;
    int rect_x, rect_y, rect_w, rect_h;
    int tiles_x, tiles_y;
    int tile_lg;           // log2 of the tile width (tile = 2^tile_lg x 64 / 2^tile_lg pixels)
    uint32_t ray_base;     // first ray slot of this job in the batch (64 slots per 8x8 tile)
    uint32_t ray_count;    // tiles_x * tiles_y * 64
    int max_steps;
    int bisect_budget;
    int refine_budget;
    int hq;
    float max_dist;
    float eps_c, eps_f;
    float footprint;       // hq_footprint_budget_px * 2 / resolution.y
    int shadow_on;         // light 0 casts a soft shadow into the G-buffer
    int light_type;        // 0 point, 1 global
    int shadow_steps;
    int normal_mode;       // nd_normal_mode
    float light_v[3];      // normalised direction (global) or position (point)
    float light_k;         // penumbra k = mix(softness, 512, hard)
    float light_far;       // 2 * max_dist (global light distance)
    float _pad2;
    // empty-space accelerator (template.wgsl:426-447): grid [level][z][y][x] of safe skip
    // distances, or null; cascade corners and cells from the Uniforms
    const float* accel;
    int accel_res, accel_levels;
    float accel_safety, _pad3;
    float accel_prm[ND_MAX_CASCADES][4];
    float* out_depth;
    float* out_normal;
    float* out_pos;
    float* out_trap;
    uint8_t* out_hit;
    float* out_shadow;
    uint32_t* out_evals;
    float* out_color;      // fs_simple colour (3 floats per pixel), or null
    float* out_albedo;     // fs_aov albedo (3 floats per pixel), or null
};

// Device counters of one batch (zeroed by a kernel before every batch).
struct DevCounters {
    uint32_t ray_next;       // march queue head (ray slots)
    uint32_t hit_count;      // hits appended by the march pass
    uint32_t surf_next;      // surface pass head (hit indices)
    uint32_t shadow_next;    // shadow queue head (hit indices)
    unsigned long long de_march;
    unsigned long long de_surface;
    unsigned long long de_shadow;
    unsigned long long march_rounds;  // wave rounds of the persistent march (one DE slot per lane each)
    unsigned long long hits;         // hits appended (the hit list also holds empty slots)
    unsigned long long clock_start;  // wall clock (s_memrealtime) when the batch started
    uint32_t rounds_capped;  // waves that stopped at the round cap
    uint32_t _pad;
};

// Per-hit SoA state passed from the march pass to the surface and shadow passes.
struct DevHitBuffers {
    uint32_t* ray;     // global ray slot of the hit (march pass)
    uint16_t* job;     // job index (march pass)
    float* t;          // hit distance (march pass)
    uint32_t* opix;    // output pixel index inside the job's rect (surface pass)
    uint16_t* steps;   // shadow step budget, 0 = no shadow march (surface pass)
    float* sox;        // shadow ray origin (surface pass)
    float* soy;
    float* soz;
    float* sdx;        // shadow ray direction
    float* sdy;
    float* sdz;
    float* sdist;      // light distance
    uint32_t capacity;
    uint32_t _pad;
};
