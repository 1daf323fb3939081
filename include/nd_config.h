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

#ifdef __cplusplus
}
#endif

#endif /* ND_CONFIG_H */
