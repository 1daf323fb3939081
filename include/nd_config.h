/*
 * nd_config.h - host-side construction of the Uniforms blocks of the five workloads.
 *
 * This mirrors the reference's buildUniforms / stampEyeBasis (src/app/export_image.zig:
 * 90-167) and the Zig struct defaults (src/gpu/fractal_renderer.zig:85-363): every field
 * starts at the reference default, then the workload sets the camera, formula, trap,
 * light and march precision. All randomness is a seeded splitmix64 -> xoshiro256**
 * stream, so C++, Python and the oracle see identical blocks.
 *
 * Workloads:
 *   1  Mandelbulb power 8, 8 iterations, 640x360, camera (0,0,-2.5) -> origin, 60 deg fov
 *   2  Mandelbox (-1.5, .5, 1, 1, 15 it), 1920x1080, 16 seeded cameras on radius 4, global
 *      light (.577,.577,-.577), soft shadow k = 16
 *   3  four-fractal smooth-min scene for the Monte Carlo path trace, 1920x1080
 *   4  256 mixed jobs: Mandelbulb (power U[2,12]) or Mandelbox (scale U[-2.5,-1.2] u
 *      U[1.5,3]), 8..20 iterations, side log-uniform 64..1024, camera radius U[2.5,5]
 *   5  differentiable scene: Mandelbulb + Mandelbox union, 512x512, 8 views on radius 3
 */
#ifndef ND_CONFIG_H
#define ND_CONFIG_H

#include <stdint.h>

#include "nd_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ND_DEFAULT_SEED 20261203ull

/* Number of Uniforms blocks (jobs / cameras / views) of a workload, 0 if unknown. */
uint32_t nd_config_job_count(uint32_t config);

/* Fills *out with the reference-default Uniforms (fractal_renderer.zig field defaults,
 * unused instance slots zeroed with code = -1 as export_image.zig:157-159 does). */
void nd_uniforms_defaults(nd_uniforms* out);

/* Builds block `index` of workload `config` from `seed` (0 = ND_DEFAULT_SEED). */
nd_status nd_config_build(uint32_t config, uint32_t index, uint64_t seed, nd_uniforms* out);

/* Orthonormal camera basis looking from pos at target, scaled for a vertical fov in
 * degrees: right = normalize(F x up_world), up = normalize(right x F) (camera.zig:95-106),
 * both scaled by tan(fov/2) so that 2 atan(|up|) = fov. */
void nd_camera_look_at(nd_uniforms* u, const float pos[3], const float target[3], float fov_deg);

/* Epsilon of the workloads: eps(t) = fraction * pixel_angle * t with pixel_angle =
 * 2 tan(fov/2) / H and epsilon_floor = 0. Workloads 1-4 use fraction 1e-3; the
 * differentiable workload 5 uses a whole pixel footprint (fraction 1), where the hit depth
 * is a smooth function of the parameters on the scale of the finite-difference step. */
void nd_set_pixel_epsilon(nd_uniforms* u, float fov_deg, float fraction);

/* Enables the empty-space accelerator as the host stamps it (stampAccelUniforms,
 * fr.zig:3073-3080): accel_res, accel_levels (<= 8), accel_safety, and accel_params from
 * placeCascades around u->camera_pos for u->max_dist (accel.zig:34-58). The app's defaults are
 * res 64, levels 6, safety 0.92 (accel.zig:13-14, accel_state.zig:19). */
void nd_accel_setup(nd_uniforms* u, uint32_t res, uint32_t levels, float safety);

#define ND_DREAM_PATH_MAX 256

/* What a .dream scene file asks of an export beyond its Uniforms block. */
typedef struct nd_dream_info {
    uint32_t version;            /* the file's format version (at most 12) */
    int32_t export_width;        /* the saved export size */
    int32_t export_height;
    uint32_t width;              /* the size the block was built for */
    uint32_t height;
    uint32_t mc_enabled;         /* the Monte Carlo renderer is on in the file */
    float mc_export_samples;     /* samples per pixel of an export (mc.export_samples) */
    float mc_max_samples;        /* the interactive view's sample cap (mc.max_samples) */
    uint32_t photons_wanted;     /* an export would trace a photon map first (photon.enabled, not
                                    replaced by approximations, 3D view); the block's photon_enabled
                                    stays 0 (no map is built here) */
    uint32_t sky_image;          /* the sky names an image: no image is loaded, the procedural sky
                                    is used, as when the app cannot load it */
    uint32_t stereo;             /* the file's stereo export is on: the block is the mono view */
    uint32_t screen_shaders;     /* post effects in the file (not part of the block) */
    uint32_t keyframes;          /* animation keys in the file (a still export does not apply them) */
    uint32_t n_codes;            /* formulas in the scene: bind them with nd_bind_formulas */
    int32_t code_to_builtin[ND_MAX_CODES];  /* code c -> nd_builtin_formula (-1 past n_codes) */
    /* the image each formula names ("image" of a formula, scene_file.zig FormulaDto), by image
     * slot: instance * 4 for its own formula, + 1 + j for mix-in j; "" for none. The host decodes
     * each file and binds it with nd_bind_image, as the app's HeightMaps.sync does. */
    char image_path[ND_IMAGE_SLOTS][ND_DREAM_PATH_MAX];
} nd_dream_info;

/* Builds the Uniforms block an export of a .dream scene file (JSON text, len bytes) renders:
 * the file is read as the app reads it (parseScene / applyScene, src/app/scene_file.zig:
 * 447-755, 811-855: unknown fields ignored, fields without a default required, at most 4
 * instances, 3 mix-ins, 16 colour stops, 8 lights, 4 fog emitters kept), then the block is
 * filled as buildUniforms does for an export (src/app/export_image.zig:97-167) with the
 * file's render settings, plus the renderer's formula codes, approximation and sky stamps.
 * width / height 0 take the file's export size. Formulas are named by file: "" is the built-in
 * Mandelbulb, otherwise the file's stem must be one of nd_builtin_formula's (with its @param
 * table compiled in; saved values override by name). Errors: ND_E_INVALID for malformed JSON,
 * a missing required field or a wrong type; ND_E_UNSUPPORTED for a file from a newer version;
 * ND_E_INVALID naming the field for warps, particle systems, the FFT view or an unknown formula
 * file, which this library does not render. Image formulas load; their image paths come back in
 * info->image_path. err (err_cap bytes, may be null)
 * receives the message. */
nd_status nd_scene_from_dream(const char* text, size_t len, uint32_t width, uint32_t height, nd_uniforms* out,
                              nd_dream_info* info, char* err, size_t err_cap);

#ifdef __cplusplus
}
#endif

#endif /* ND_CONFIG_H */
