"""ctypes mirror of include/nd_uniforms.h and include/nd_hip.h.

The Uniforms block is the reference's per-sample parameter block
(src/gpu/fractal_renderer.zig:85-363): 12,016 bytes of 32-bit floats in declaration order.
"""
import ctypes as C

F = C.c_float


def _f3():
    return F * 3


class ColorStop(C.Structure):
    _fields_ = [("color", _f3()), ("position", F), ("glossiness", F), ("transparency", F),
                ("reflectiveness", F), ("ior", F), ("subsurface", F), ("abbe", F),
                ("inner_max_steps", F), ("roughness", F), ("film_thickness", F), ("film_ior", F),
                ("film_strength", F), ("film_angle_scale", F), ("film_perturb", F),
                ("film_perturb_scale", F), ("subsurface_radius", F), ("_pad_film1", F)]


class MixinParams(C.Structure):
    _fields_ = [("params", F * 16), ("iterations", F), ("code", F), ("formula_key", F), ("_pad0", F)]


class FractalInstance(C.Structure):
    _fields_ = [("offset", _f3()), ("blend_k", F), ("combine_mode", F), ("color_count", F),
                ("step_safety", F), ("hidden", F), ("scale", _f3()), ("code", F), ("rotation", _f3()),
                ("formula_key", F), ("params", F * 16), ("mixin_count", F), ("hybrid_base_iters", F),
                ("hybrid_total_iters", F), ("trap_repeat", F), ("trap_center", _f3()), ("trap_shape", F),
                ("trap_box", _f3()), ("trap_mode", F), ("trap_radius", F), ("trap_tube", F),
                ("trap_span", F), ("trap_offset", F), ("mixins", MixinParams * 3),
                ("colors", ColorStop * 16), ("frame", (F * 4) * 3)]


class Light(C.Structure):
    _fields_ = [("color", _f3()), ("brightness", F), ("position_or_direction", _f3()), ("light_type", F),
                ("shadow_softness", F), ("cast_shadows", F), ("hard_shadows", F), ("spread", F),
                ("ray_direction", _f3()), ("waist", F)]


class FogEmitter(C.Structure):
    _fields_ = [("position", _f3()), ("radius", F), ("color", _f3()), ("density", F), ("softness", F),
                ("anisotropy", F), ("_pad0", F), ("_pad1", F)]


class Warp(C.Structure):
    _fields_ = [("center", _f3()), ("region_kind", F), ("extent", _f3()), ("falloff", F),
                ("rotation", _f3()), ("sub_kind", F), ("params0", F * 4), ("params1", F * 4),
                ("strength", F), ("lip_mult", F), ("lip_grad", F), ("safety", F), ("frame", (F * 4) * 3)]


class ParticleSystem(C.Structure):
    _fields_ = [("center", _f3()), ("mode", F), ("dot_style", F), ("glow", F), ("combine_mode", F),
                ("blend_k", F), ("color_count", F), ("generation", F), ("glow_extent", F), ("_pad0", F),
                ("colors", ColorStop * 16)]


class Sky(C.Structure):
    _fields_ = [("mode", F), ("intensity", F), ("falloff", F), ("photons", F), ("zenith", _f3()),
                ("sun_cos", F), ("horizon", _f3()), ("sun_intensity", F), ("ground", _f3()), ("yaw", F),
                ("sun_color", _f3()), ("_pad0", F), ("sun_dir", _f3()), ("_pad1", F)]


_U_FIELDS = [
    ("camera_pos", _f3()), ("time", F), ("camera_right", _f3()), ("max_steps", F), ("camera_up", _f3()),
    ("max_dist", F), ("camera_forward", _f3()), ("instance_count", F), ("resolution", F * 2),
    ("light_count", F), ("max_reflection_bounces", F), ("high_quality", F), ("epsilon_coefficient", F),
    ("epsilon_floor", F), ("hq_footprint_budget_px", F), ("tile_scale", F * 2), ("tile_bias", F * 2),
    ("dof_enabled", F), ("focus_distance", F), ("aperture", F), ("focus_range", F), ("refine_fast", F),
    ("refine_hq", F), ("mc_enabled", F), ("mc_sample", F), ("instances", FractalInstance * 4),
    ("lights", Light * 8), ("fog_count", F), ("fog_samples", F), ("mode_2d", F), ("slice_zoom", F),
    ("fog_emitters", FogEmitter * 4), ("light_bounces", F), ("accel_enabled", F), ("accel_levels", F),
    ("accel_res", F), ("accel_safety", F), ("accel_z_offset", F), ("fft_active", F),
    ("fft_target_slot", F), ("accel_params", (F * 4) * 8), ("warp_count", F), ("has_transparency", F),
    ("photon_prefilter", F), ("march_verify", F), ("warps", Warp * 4), ("photon_enabled", F),
    ("photon_radius", F), ("photon_cell", F), ("photon_table", F), ("photon_paths", F),
    ("photon_seed", F), ("photon_intensity", F), ("photon_path_offset", F), ("photon_centre", _f3()),
    ("photon_extent", F), ("sky", Sky), ("photon_pass", F), ("photon_pool", F),
    ("photon_cap_surface", F), ("photon_fixed_unit", F), ("photon_volume_scale", F),
    ("photon_hash_salt", F), ("photon_dispersion_soft", F), ("debug_parts", F), ("fft_axis", F),
    ("fft_box_radius", F), ("fft_cloud_density", F), ("fft_lowpass", F), ("fft_highpass", F),
    ("fft_normalize", F), ("photon_bounce_scale", F), ("photon_aim", F), ("carve_count", F),
    ("hotspot_scale", F), ("hotspot_sources", F), ("overstep_scale", F), ("carves", (F * 4) * 16),
    ("particle_systems", ParticleSystem * 2)]
_U_FIELDS += [(n, F) for n in (
    "approx_flags approx_caustic_strength approx_caustic_scale approx_bounce approx_tint "
    "approx_lod_start approx_lod_strength mc_light_pick adapt_enabled adapt_threshold "
    "adapt_min_samples adapt_row_width flat_coloring flat_interior flat_bailout flat_extra_iters "
    "flat_band_density flat_band_offset flat_relief flat_light_angle flat_contours "
    "flat_contour_density flat_slab_slices flat_slab_depth flat_slab_opacity flat_samples "
    "flat_halo mc_inner_samples fog_tracking primary_cache_key fog_vis_enabled fog_vis_layers").split()]
_U_FIELDS += [("fog_vis_origin", _f3()), ("fog_vis_cell", F), ("fog_vis_dims", _f3()),
              ("fog_vis_z_offset", F), ("fog_vis_light_layer", (F * 4) * 2),
              ("fog_vis_layer_light", (F * 4) * 2)]


class Uniforms(C.Structure):
    _fields_ = _U_FIELDS


UNIFORMS_SIZE = 12016
assert C.sizeof(Uniforms) == UNIFORMS_SIZE, C.sizeof(Uniforms)

# C ABI records (include/nd_hip.h)
U32 = C.c_uint32


class MarchOpts(C.Structure):
    _fields_ = [("normal_mode", U32), ("shadow_max_steps", U32), ("shadows", U32), ("kernel", U32)]


class Rect(C.Structure):
    _fields_ = [("x", U32), ("y", U32), ("w", U32), ("h", U32)]


class GBufferOut(C.Structure):
    _fields_ = [("depth", C.c_void_p), ("normal", C.c_void_p), ("position", C.c_void_p), ("trap", C.c_void_p),
                ("hit", C.c_void_p), ("shadow", C.c_void_p), ("evals", C.c_void_p), ("color", C.c_void_p),
                ("albedo", C.c_void_p), ("device_ptrs", U32), ("_pad", U32)]


class Job(C.Structure):
    _fields_ = [("uniforms", C.POINTER(Uniforms)), ("rect", Rect), ("out", GBufferOut)]


class BatchStats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("hits", C.c_uint64), ("de_evals", C.c_uint64),
                ("de_evals_march", C.c_uint64), ("de_evals_surface", C.c_uint64),
                ("de_evals_shadow", C.c_uint64), ("ms_total", F), ("ms_march", F), ("ms_surface", F),
                ("ms_shadow", F), ("kernels", U32), ("graph", U32), ("march_wave_rounds", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("waves_per_cu", U32), ("refill_min_idle", U32), ("chunk_rays", U32), ("flags", U32),
                ("round_cap", U32), ("tile_width", U32), ("refill_period", U32),
                ("mc_pool_log2", U32)]


NOALLOC_MAX_RAYS = 1 << 19
NOALLOC_MAX_GRID = 4096

TUNE_TRACK_LATENCY = 1
TUNE_NO_GRAPH = 2
TUNE_GBUFFER_GRAPH = 4


class AdjointIO(C.Structure):
    _fields_ = [("target_depth", C.c_void_p), ("depth_out", C.c_void_p), ("dtdp_out", C.c_void_p),
                ("grad_out", C.c_void_p), ("loss_out", C.c_void_p), ("pixels_used", C.c_void_p),
                ("device_ptrs", U32), ("_pad", U32)]


class ImageOut(C.Structure):
    _fields_ = [("color", C.c_void_p), ("depth", C.c_void_p), ("normal", C.c_void_p),
                ("device_ptrs", U32), ("_pad", U32)]


class TraceStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("segments", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("de_evals", C.c_uint64), ("ms_total", F), ("max_depth_reached", U32), ("_pad", U32)]


# status codes
ND_OK, ND_E_INVALID, ND_E_UNSUPPORTED_FORMULA, ND_E_OOM, ND_E_DEVICE, ND_E_TIMEOUT, ND_E_CANCELLED, \
    ND_E_UNSUPPORTED = range(8)
STATUS_NAMES = {0: "ND_OK", 1: "ND_E_INVALID", 2: "ND_E_UNSUPPORTED_FORMULA", 3: "ND_E_OOM",
                4: "ND_E_DEVICE", 5: "ND_E_TIMEOUT", 6: "ND_E_CANCELLED", 7: "ND_E_UNSUPPORTED"}

NORMAL_ANALYTIC, NORMAL_TETRA, NORMAL_CENTRAL = 0, 1, 2
KERNEL_AUTO, KERNEL_PLAIN, KERNEL_PERSISTENT = 0, 1, 2
FORMULA_MANDELBULB, FORMULA_MANDELBOX, FORMULA_MENGER3, FORMULA_KLEINIAN = 0, 1, 2, 3
# nd_builtin_formula id -> the formula file it implements (include/nd_hip.h); a host maps
# each entry of its formula code table to an id by the file it was loaded from.
FORMULA_FILES = [
    "src/gpu/shaders/mandelbrot.wgsl", "formulas/amazingbox_box.wgsl", "formulas/menger3_ifs.wgsl",
    "formulas/kleinian_kleinian.wgsl"] + ["formulas/%s.wgsl" % f for f in (
        "sphere_primitive", "cube_primitive", "cylinder_primitive", "torus_primitive", "pyramid_primitive",
        "prism_primitive", "sierpinski3_ifs", "cubejulia_ifs_julia", "juliakifs_ifs_julia", "diagonaltetra_ifs_fold",
        "starkifs_ifs_fold", "kaleidojulia_ifs_fold_julia", "bulbtetra_hybrid_bulb_ifs", "icosahedralifs_ifs",
        "icosahedraljulia_ifs_julia", "icosahedraltwist_ifs_twist", "diagonalicosa_ifs_fold", "hybridkifs_hybrid_ifs",
        "kaleidoscopicifs_ifs_twist", "juliabox_box_julia", "amazingsurf_box_fold", "mandelboxfoam_box_fold",
        "mandelboxsort_box_fold", "mandelboxdiag_box_fold", "mandelboxinv_box_fold", "mandelboxtwist_box_twist",
        "prismkifs_box_fold", "buffalo_box", "buffalojulia_box_julia", "buffalotwist_box_twist",
        "apollonian_kleinian", "apollonianjulia_kleinian_julia", "kleinianjulia_kleinian_julia",
        "pseudokleinian_kleinian", "juliabulb_bulb_julia", "mandelbulbmirror_bulb_fold", "mandelbulbtwist_bulb_twist",
        "mandelbulbdual_hybrid_bulb", "quatjulia_bulb_julia", "bulbbox_hybrid_bulb_box", "mandalay_box_fold",
        "wobblekifs_ifs_twist_julia", "radioactive_ifs_carve_twist", "recursivesphere_ifs_carve_twist", "tree_ifs",
        "kleinianlimit_kleinian", "fbmsdf_primitive_carve", "seascape_primitive", "dragon_ifs_twist",
        "romanesco_ifs", "heightmap_primitive")]
FORMULA_COUNT = len(FORMULA_FILES)
PARAM_COUNT = 6

# every symbol the headers declare (include/nd_hip.h, include/nd_config.h)
class RadianceIO(C.Structure):
    _fields_ = [("target_rgb", C.c_void_p), ("rgb_out", C.c_void_p), ("drgb_out", C.c_void_p),
                ("grad_out", C.c_void_p), ("loss_out", C.c_void_p), ("pixels_used", C.c_void_p),
                ("device_ptrs", U32), ("_pad", U32)]


class MeshGrid(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("cell", C.c_float), ("iso", C.c_float), ("span", C.c_uint32),
                ("local_offset", C.c_uint32), ("device_ptrs", C.c_uint32)]


EXPORTS = ["nd_ctx_create", "nd_ctx_destroy", "nd_last_error", "nd_sync", "nd_set_cancel", "nd_ctx_stream",
           "nd_bind_formulas", "nd_render_gbuffer", "nd_render_gbuffer_noalloc", "nd_render_batch", "nd_batch_job_latency",
           "nd_get_tuning", "nd_set_tuning", "nd_autotune", "nd_adjoint", "nd_path_trace", "nd_mesh_points", "nd_mesh_probe", "nd_accel_probe", "nd_accel_setup",
           "nd_config_job_count", "nd_uniforms_defaults", "nd_config_build", "nd_camera_look_at",
           "nd_set_pixel_epsilon", "nd_scene_from_dream", "nd_radiance_adjoint", "nd_bind_image", "nd_image_pack"]

# height-map images (nd_bind_image / nd_image_pack)
IMAGE_RGBA16 = 0
IMAGE_RGBA32F = 1
IMAGE_SLOTS = 16
IMAGE_MAX_EDGE = 2048


class ImageDesc(C.Structure):
    _fields_ = [("width", U32), ("height", U32), ("words", U32), ("slope", F * 4)]


class DreamInfo(C.Structure):
    _fields_ = [("version", U32), ("export_width", C.c_int32), ("export_height", C.c_int32), ("width", U32),
                ("height", U32), ("mc_enabled", U32), ("mc_export_samples", F), ("mc_max_samples", F),
                ("photons_wanted", U32), ("sky_image", U32), ("stereo", U32), ("screen_shaders", U32),
                ("keyframes", U32), ("n_codes", U32), ("code_to_builtin", C.c_int32 * 16),
                ("image_path", (C.c_char * 256) * 16)]
