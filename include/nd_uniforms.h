/*
 * nd_uniforms.h - byte-exact C mirror of the renderer's per-sample parameter block.
 *
 * The host writes one `nd_uniforms` (12,016 bytes) per sample; the march, G-buffer,
 * path-trace and adjoint entry points read it unchanged. Every field is a 32-bit float
 * (enums and counts are stored as floats, exactly as the reference stores them), so the
 * C layout is the plain declaration order with no padding.
 *
 * Reference: src/gpu/fractal_renderer.zig:85-363 (Zig extern structs, the CPU side of
 * the ABI) and src/gpu/shaders/template.wgsl:1-315, 405-420 (the WGSL side). The
 * layout asserts at fractal_renderer.zig:106-110, 151-157, 199-201, 217-220, 365-390 are
 * restated as static asserts at the end of this file.
 *
 * Plain C99/C11, no C++ types: consumable by Zig's translate-c like the webgpu shim.
 */
#ifndef ND_UNIFORMS_H
#define ND_UNIFORMS_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ND_MAX_INSTANCES 4
#define ND_MAX_MIXINS 3
#define ND_MAX_COLOR_STOPS 16
#define ND_MAX_PARAMS 16
#define ND_MAX_LIGHTS 8
#define ND_MAX_FOG_EMITTERS 4
#define ND_MAX_WARPS 4
#define ND_MAX_CASCADES 8
#define ND_MAX_CARVES 16
#define ND_MAX_PARTICLE_SYSTEMS 2
#define ND_NO_CODE (-1.0f)

/* Material stop of an orbit-trap colour strip (fractal_renderer.zig:85-104). */
typedef struct nd_color_stop {
    float color[3];
    float position;
    float glossiness;
    float transparency;
    float reflectiveness;
    float ior;
    float subsurface;
    float abbe;
    float inner_max_steps;
    float roughness;
    float film_thickness;
    float film_ior;
    float film_strength;
    float film_angle_scale;
    float film_perturb;
    float film_perturb_scale;
    float subsurface_radius;
    float _pad_film1;
} nd_color_stop;

/* Hybrid mix-in formula slot (fractal_renderer.zig:112-117). */
typedef struct nd_mixin_params {
    float params[ND_MAX_PARAMS];
    float iterations;
    float code;
    float formula_key;
    float _pad0;
} nd_mixin_params;

/* One fractal instance (fractal_renderer.zig:119-145). `frame` is the rotation matrix,
 * column-major, each column padded to 4 floats (camera.zig:41-60). */
typedef struct nd_fractal_instance {
    float offset[3];
    float blend_k;
    float combine_mode;
    float color_count;
    float step_safety;
    float hidden;
    float scale[3];
    float code;
    float rotation[3];
    float formula_key;
    float params[ND_MAX_PARAMS];
    float mixin_count;
    float hybrid_base_iters;
    float hybrid_total_iters;
    float trap_repeat;
    float trap_center[3];
    float trap_shape;
    float trap_box[3];
    float trap_mode;
    float trap_radius;
    float trap_tube;
    float trap_span;
    float trap_offset;
    nd_mixin_params mixins[ND_MAX_MIXINS];
    nd_color_stop colors[ND_MAX_COLOR_STOPS];
    float frame[3][4];
} nd_fractal_instance;

/* Light (fractal_renderer.zig:159-170). light_type: 0 point, 1 global (directional), 2 ray. */
typedef struct nd_light {
    float color[3];
    float brightness;
    float position_or_direction[3];
    float light_type;
    float shadow_softness;
    float cast_shadows;
    float hard_shadows;
    float spread;
    float ray_direction[3];
    float waist;
} nd_light;

typedef struct nd_fog_emitter {
    float position[3];
    float radius;
    float color[3];
    float density;
    float softness;
    float anisotropy;
    float _pad0;
    float _pad1;
} nd_fog_emitter;

typedef struct nd_warp {
    float center[3];
    float region_kind;
    float extent[3];
    float falloff;
    float rotation[3];
    float sub_kind;
    float params0[4];
    float params1[4];
    float strength;
    float lip_mult;
    float lip_grad;
    float safety;
    float frame[3][4];
} nd_warp;

typedef struct nd_particle_system {
    float center[3];
    float mode;
    float dot_style;
    float glow;
    float combine_mode;
    float blend_k;
    float color_count;
    float generation;
    float glow_extent;
    float _pad0;
    nd_color_stop colors[ND_MAX_COLOR_STOPS];
} nd_particle_system;

/* Procedural sky (fractal_renderer.zig:348-363, template.wgsl:405-420). */
typedef struct nd_sky {
    float mode;
    float intensity;
    float falloff;
    float photons;
    float zenith[3];
    float sun_cos;
    float horizon[3];
    float sun_intensity;
    float ground[3];
    float yaw;
    float sun_color[3];
    float _pad0;
    float sun_dir[3];
    float _pad1;
} nd_sky;

/* The per-sample block (fractal_renderer.zig:228-346). */
typedef struct nd_uniforms {
    float camera_pos[3];
    float time;
    float camera_right[3];
    float max_steps;
    float camera_up[3];
    float max_dist;
    float camera_forward[3];
    float instance_count;
    float resolution[2];
    float light_count;
    float max_reflection_bounces;
    float high_quality;
    float epsilon_coefficient;
    float epsilon_floor;
    float hq_footprint_budget_px;
    float tile_scale[2];
    float tile_bias[2];
    float dof_enabled;
    float focus_distance;
    float aperture;
    float focus_range;
    float refine_fast;
    float refine_hq;
    float mc_enabled;
    float mc_sample;
    nd_fractal_instance instances[ND_MAX_INSTANCES];
    nd_light lights[ND_MAX_LIGHTS];
    float fog_count;
    float fog_samples;
    float mode_2d;
    float slice_zoom;
    nd_fog_emitter fog_emitters[ND_MAX_FOG_EMITTERS];
    float light_bounces;
    float accel_enabled;
    float accel_levels;
    float accel_res;
    float accel_safety;
    float accel_z_offset;
    float fft_active;
    float fft_target_slot;
    float accel_params[ND_MAX_CASCADES][4];
    float warp_count;
    float has_transparency;
    float photon_prefilter;
    float march_verify;
    nd_warp warps[ND_MAX_WARPS];
    float photon_enabled;
    float photon_radius;
    float photon_cell;
    float photon_table;
    float photon_paths;
    float photon_seed;
    float photon_intensity;
    float photon_path_offset;
    float photon_centre[3];
    float photon_extent;
    nd_sky sky;
    float photon_pass;
    float photon_pool;
    float photon_cap_surface;
    float photon_fixed_unit;
    float photon_volume_scale;
    float photon_hash_salt;
    float photon_dispersion_soft;
    float debug_parts;
    float fft_axis;
    float fft_box_radius;
    float fft_cloud_density;
    float fft_lowpass;
    float fft_highpass;
    float fft_normalize;
    float photon_bounce_scale;
    float photon_aim;
    float carve_count;
    float hotspot_scale;
    float hotspot_sources;
    float overstep_scale;
    float carves[ND_MAX_CARVES][4];
    nd_particle_system particle_systems[ND_MAX_PARTICLE_SYSTEMS];
    float approx_flags;
    float approx_caustic_strength;
    float approx_caustic_scale;
    float approx_bounce;
    float approx_tint;
    float approx_lod_start;
    float approx_lod_strength;
    float mc_light_pick;
    float adapt_enabled;
    float adapt_threshold;
    float adapt_min_samples;
    float adapt_row_width;
    float flat_coloring;
    float flat_interior;
    float flat_bailout;
    float flat_extra_iters;
    float flat_band_density;
    float flat_band_offset;
    float flat_relief;
    float flat_light_angle;
    float flat_contours;
    float flat_contour_density;
    float flat_slab_slices;
    float flat_slab_depth;
    float flat_slab_opacity;
    float flat_samples;
    float flat_halo;
    float mc_inner_samples;
    float fog_tracking;
    float primary_cache_key;
    float fog_vis_enabled;
    float fog_vis_layers;
    float fog_vis_origin[3];
    float fog_vis_cell;
    float fog_vis_dims[3];
    float fog_vis_z_offset;
    float fog_vis_light_layer[ND_MAX_LIGHTS / 4][4];
    float fog_vis_layer_light[ND_MAX_LIGHTS / 4][4];
} nd_uniforms;

#define ND_UNIFORMS_SIZE 12016u

/* debug_parts bits (template.wgsl:333-350). A set bit switches a part OFF. */
#define ND_PART_COLOR_STRIPS (1u << 0)
#define ND_PART_DIRECT_LIGHTS (1u << 1)
#define ND_PART_SPECULAR (1u << 2)
#define ND_PART_SHADOWS (1u << 3)
#define ND_PART_AMBIENT_OCCLUSION (1u << 4)
#define ND_PART_FOG (1u << 5)
#define ND_PART_REFLECTIONS (1u << 6)
#define ND_PART_REFRACTION (1u << 7)
#define ND_PART_PHOTON_MAP (1u << 8)
#define ND_PART_SUBSURFACE (1u << 9)
#define ND_PART_IRIDESCENCE (1u << 10)
#define ND_PART_DISPERSION (1u << 11)
#define ND_PART_SKY (1u << 12)
#define ND_PART_DEPTH_OF_FIELD (1u << 13)
#define ND_PART_MC_INDIRECT (1u << 14)
#define ND_PART_GEOMETRY_ONLY (1u << 16)

/* Layout asserts: the reference's own (fractal_renderer.zig) plus the offsets the HIP
 * kernels read. */
#if defined(__cplusplus)
#define ND_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define ND_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif
ND_STATIC_ASSERT(sizeof(nd_color_stop) == 80, "ColorStop is 80 B");
ND_STATIC_ASSERT(offsetof(nd_color_stop, roughness) == 44, "ColorStop.roughness @44");
ND_STATIC_ASSERT(offsetof(nd_color_stop, film_thickness) == 48, "ColorStop.film_thickness @48");
ND_STATIC_ASSERT(sizeof(nd_mixin_params) == 80, "MixinParams is 80 B");
ND_STATIC_ASSERT(offsetof(nd_fractal_instance, trap_center) % 16 == 0, "trap_center 16-aligned");
ND_STATIC_ASSERT(offsetof(nd_fractal_instance, trap_box) % 16 == 0, "trap_box 16-aligned");
ND_STATIC_ASSERT(offsetof(nd_fractal_instance, mixins) % 16 == 0, "mixins 16-aligned");
ND_STATIC_ASSERT(offsetof(nd_fractal_instance, frame) % 16 == 0, "frame 16-aligned");
ND_STATIC_ASSERT(offsetof(nd_fractal_instance, params) == 64, "params @64");
ND_STATIC_ASSERT(offsetof(nd_fractal_instance, code) == 44, "code @44");
ND_STATIC_ASSERT(offsetof(nd_fractal_instance, colors) == 432, "colors @432");
ND_STATIC_ASSERT(offsetof(nd_fractal_instance, frame) == 1712, "frame @1712");
ND_STATIC_ASSERT(sizeof(nd_fractal_instance) == 1760, "FractalInstance is 1760 B");
ND_STATIC_ASSERT(sizeof(nd_light) == 64, "Light is 64 B");
ND_STATIC_ASSERT(offsetof(nd_warp, frame) == 96, "Warp.frame @96");
ND_STATIC_ASSERT(sizeof(nd_warp) == 144, "Warp is 144 B");
ND_STATIC_ASSERT(sizeof(nd_particle_system) == 48 + 80 * ND_MAX_COLOR_STOPS, "ParticleSystem size");
ND_STATIC_ASSERT(offsetof(nd_particle_system, colors) == 48, "ParticleSystem.colors @48");
ND_STATIC_ASSERT(sizeof(nd_sky) == 96, "Sky is 96 B");
ND_STATIC_ASSERT(offsetof(nd_sky, zenith) % 16 == 0 && offsetof(nd_sky, horizon) % 16 == 0 &&
                 offsetof(nd_sky, ground) % 16 == 0 && offsetof(nd_sky, sun_color) % 16 == 0 &&
                 offsetof(nd_sky, sun_dir) % 16 == 0, "Sky vec3 fields 16-aligned");
ND_STATIC_ASSERT(sizeof(nd_uniforms) == ND_UNIFORMS_SIZE, "Uniforms is 12016 B");
ND_STATIC_ASSERT(sizeof(nd_uniforms) % 16 == 0, "Uniforms size 16-aligned");
ND_STATIC_ASSERT(offsetof(nd_uniforms, instances) == 144, "instances @144");
ND_STATIC_ASSERT(offsetof(nd_uniforms, lights) == 7184, "lights @7184");
ND_STATIC_ASSERT(offsetof(nd_uniforms, accel_params) == 7936, "accel_params @7936");
ND_STATIC_ASSERT(offsetof(nd_uniforms, march_verify) == 8076, "march_verify @8076");
ND_STATIC_ASSERT(offsetof(nd_uniforms, warps) == 8080, "warps @8080");
ND_STATIC_ASSERT(offsetof(nd_uniforms, photon_enabled) == 8656, "photon_enabled @8656");
ND_STATIC_ASSERT(offsetof(nd_uniforms, photon_centre) % 16 == 0, "photon_centre 16-aligned");
ND_STATIC_ASSERT(offsetof(nd_uniforms, sky) == 8704, "sky @8704");
ND_STATIC_ASSERT(offsetof(nd_uniforms, photon_pass) % 16 == 0, "photon_pass 16-aligned");
ND_STATIC_ASSERT(offsetof(nd_uniforms, debug_parts) == 8828, "debug_parts @8828");
ND_STATIC_ASSERT(offsetof(nd_uniforms, carves) == 8880, "carves @8880");
ND_STATIC_ASSERT(offsetof(nd_uniforms, particle_systems) == 9136, "particle_systems @9136");
ND_STATIC_ASSERT(offsetof(nd_uniforms, approx_flags) == 11792, "approx_flags @11792");
ND_STATIC_ASSERT(offsetof(nd_uniforms, adapt_enabled) == 11824, "adapt_enabled @11824");
ND_STATIC_ASSERT(offsetof(nd_uniforms, flat_coloring) % 16 == 0, "flat_coloring 16-aligned");
ND_STATIC_ASSERT(offsetof(nd_uniforms, fog_tracking) % 16 == 0, "fog_tracking 16-aligned");
ND_STATIC_ASSERT(offsetof(nd_uniforms, primary_cache_key) == 11908, "primary_cache_key @11908");
ND_STATIC_ASSERT(offsetof(nd_uniforms, fog_vis_origin) % 16 == 0, "fog_vis_origin 16-aligned");
ND_STATIC_ASSERT(offsetof(nd_uniforms, fog_vis_dims) % 16 == 0, "fog_vis_dims 16-aligned");
ND_STATIC_ASSERT(offsetof(nd_uniforms, fog_vis_light_layer) == 11952, "fog_vis_light_layer @11952");

#ifdef __cplusplus
}
#endif

#endif /* ND_UNIFORMS_H */
