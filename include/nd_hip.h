/*
 * nd_hip.h - C ABI of the MI355X (gfx950) distance-estimator ray-march and path-trace
 * passes. Plain C99: no C++ types, no HIP types, plain pointers and sizes, so the Zig host
 * can bind it with translate-c exactly as it binds the webgpu shim (build.zig:364-366,
 * 394-400).
 *
 * What each entry point replaces in the reference (all paths relative to the reference
 * tree; fr.zig = src/gpu/fractal_renderer.zig):
 *
 *   nd_ctx_create / nd_ctx_destroy  Context.init + FractalRenderer.init
 *                                   (src/gpu/webgpu_context.zig:243-283, fr.zig:1769-)
 *   nd_last_error                   the uncaptured-error sink, 512-byte message buffer
 *                                   (src/gpu/webgpu_context.zig:131-146, 203-225)
 *   nd_bind_formulas                the formula code table that stampFormulaCodes writes
 *                                   into FractalInstance.code (fr.zig:602-615, 2882-2896);
 *                                   here a code selects a built-in HIP DE instead of a
 *                                   spliced WGSL function (fr.zig:490-539)
 *   nd_render_gbuffer               renderToImage/renderTile -> drawSlot(.simple) ->
 *                                   fs_simple/simple_geometry (fr.zig:3548, 4292-4298;
 *                                   template.wgsl:4492-4538) plus the light-0 soft shadow
 *                                   of shade_local (template.wgsl:2365-2375)
 *   nd_render_batch                 SampleSet.per_sample + ExportBatch (fr.zig:762-783,
 *                                   1066-1094): many frames/jobs, here one persistent launch
 *   nd_path_trace                   encodeWavefront + fs_wf_resolve (fr.zig:4334-4381;
 *                                   template.wgsl:4898-5407)
 *   nd_adjoint                      no counterpart in the reference: gradient of hit depth
 *                                   with respect to formula parameters and camera position
 *   nd_sync / nd_set_cancel         retireOldest / export cancel flag (fr.zig:1100-1137,
 *                                   src/app/export_image.zig:46-47)
 *
 * Threading: one context owns one HIP stream; calls on one context must be serialised by
 * the caller (the reference submits from the main thread only). Contexts are independent.
 * Errors: every call returns an nd_status; the message of the last failure is kept in the
 * context (nd_last_error). Nothing throws or longjmps across this boundary.
 */
#ifndef ND_HIP_H
#define ND_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "nd_uniforms.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nd_ctx nd_ctx;

typedef enum nd_status {
    ND_OK = 0,
    ND_E_INVALID = 1,              /* bad argument: size, null, range */
    ND_E_UNSUPPORTED_FORMULA = 2,  /* an instance code has no HIP DE bound; caller may fall back */
    ND_E_OOM = 3,                  /* device allocation failed */
    ND_E_DEVICE = 4,               /* HIP runtime error or no device */
    ND_E_TIMEOUT = 5,              /* a persistent kernel hit its round cap */
    ND_E_CANCELLED = 6,            /* the cancel flag was raised */
    ND_E_UNSUPPORTED = 7           /* a Uniforms feature outside this library (accel, warps, fog...) */
} nd_status;

/* Built-in distance estimators: one per formula plug-in (a formulas/*.wgsl file, or the
 * built-in mandelbrot.wgsl), with the @param slots of that file. 0-3 are the tuned ones. */
typedef enum nd_builtin_formula {
    ND_FORMULA_NONE = -1,
    ND_FORMULA_MANDELBULB = 0,  /* src/gpu/shaders/mandelbrot.wgsl */
    ND_FORMULA_MANDELBOX = 1,   /* formulas/amazingbox_box.wgsl */
    ND_FORMULA_MENGER3 = 2,     /* formulas/menger3_ifs.wgsl */
    ND_FORMULA_KLEINIAN = 3,    /* formulas/kleinian_kleinian.wgsl */
    ND_FORMULA_SPHERE = 4,             /* formulas/sphere_primitive.wgsl */
    ND_FORMULA_CUBE = 5,               /* formulas/cube_primitive.wgsl */
    ND_FORMULA_CYLINDER = 6,           /* formulas/cylinder_primitive.wgsl */
    ND_FORMULA_TORUS = 7,              /* formulas/torus_primitive.wgsl */
    ND_FORMULA_PYRAMID = 8,            /* formulas/pyramid_primitive.wgsl */
    ND_FORMULA_PRISM = 9,              /* formulas/prism_primitive.wgsl */
    ND_FORMULA_SIERPINSKI3 = 10,       /* formulas/sierpinski3_ifs.wgsl */
    ND_FORMULA_CUBEJULIA = 11,         /* formulas/cubejulia_ifs_julia.wgsl */
    ND_FORMULA_JULIAKIFS = 12,         /* formulas/juliakifs_ifs_julia.wgsl */
    ND_FORMULA_DIAGONALTETRA = 13,     /* formulas/diagonaltetra_ifs_fold.wgsl */
    ND_FORMULA_STARKIFS = 14,          /* formulas/starkifs_ifs_fold.wgsl */
    ND_FORMULA_KALEIDOJULIA = 15,      /* formulas/kaleidojulia_ifs_fold_julia.wgsl */
    ND_FORMULA_BULBTETRA = 16,         /* formulas/bulbtetra_hybrid_bulb_ifs.wgsl */
    ND_FORMULA_ICOSAHEDRALIFS = 17,    /* formulas/icosahedralifs_ifs.wgsl */
    ND_FORMULA_ICOSAHEDRALJULIA = 18,  /* formulas/icosahedraljulia_ifs_julia.wgsl */
    ND_FORMULA_ICOSAHEDRALTWIST = 19,  /* formulas/icosahedraltwist_ifs_twist.wgsl */
    ND_FORMULA_DIAGONALICOSA = 20,     /* formulas/diagonalicosa_ifs_fold.wgsl */
    ND_FORMULA_HYBRIDKIFS = 21,        /* formulas/hybridkifs_hybrid_ifs.wgsl */
    ND_FORMULA_KALEIDOSCOPICIFS = 22,  /* formulas/kaleidoscopicifs_ifs_twist.wgsl */
    ND_FORMULA_JULIABOX = 23,          /* formulas/juliabox_box_julia.wgsl */
    ND_FORMULA_AMAZINGSURF = 24,       /* formulas/amazingsurf_box_fold.wgsl */
    ND_FORMULA_MANDELBOXFOAM = 25,     /* formulas/mandelboxfoam_box_fold.wgsl */
    ND_FORMULA_MANDELBOXSORT = 26,     /* formulas/mandelboxsort_box_fold.wgsl */
    ND_FORMULA_MANDELBOXDIAG = 27,     /* formulas/mandelboxdiag_box_fold.wgsl */
    ND_FORMULA_MANDELBOXINV = 28,      /* formulas/mandelboxinv_box_fold.wgsl */
    ND_FORMULA_MANDELBOXTWIST = 29,    /* formulas/mandelboxtwist_box_twist.wgsl */
    ND_FORMULA_PRISMKIFS = 30,         /* formulas/prismkifs_box_fold.wgsl */
    ND_FORMULA_BUFFALO = 31,           /* formulas/buffalo_box.wgsl */
    ND_FORMULA_BUFFALOJULIA = 32,      /* formulas/buffalojulia_box_julia.wgsl */
    ND_FORMULA_BUFFALOTWIST = 33,      /* formulas/buffalotwist_box_twist.wgsl */
    ND_FORMULA_APOLLONIAN = 34,        /* formulas/apollonian_kleinian.wgsl */
    ND_FORMULA_APOLLONIANJULIA = 35,   /* formulas/apollonianjulia_kleinian_julia.wgsl */
    ND_FORMULA_KLEINIANJULIA = 36,     /* formulas/kleinianjulia_kleinian_julia.wgsl */
    ND_FORMULA_PSEUDOKLEINIAN = 37,    /* formulas/pseudokleinian_kleinian.wgsl */
    ND_FORMULA_JULIABULB = 38,         /* formulas/juliabulb_bulb_julia.wgsl */
    ND_FORMULA_MANDELBULBMIRROR = 39,  /* formulas/mandelbulbmirror_bulb_fold.wgsl */
    ND_FORMULA_MANDELBULBTWIST = 40,   /* formulas/mandelbulbtwist_bulb_twist.wgsl */
    ND_FORMULA_MANDELBULBDUAL = 41,    /* formulas/mandelbulbdual_hybrid_bulb.wgsl */
    ND_FORMULA_QUATJULIA = 42,         /* formulas/quatjulia_bulb_julia.wgsl */
    ND_FORMULA_BULBBOX = 43,           /* formulas/bulbbox_hybrid_bulb_box.wgsl */
    ND_FORMULA_MANDALAY = 44,          /* formulas/mandalay_box_fold.wgsl */
    ND_FORMULA_WOBBLEKIFS = 45,        /* formulas/wobblekifs_ifs_twist_julia.wgsl */
    ND_FORMULA_RADIOACTIVE = 46,       /* formulas/radioactive_ifs_carve_twist.wgsl */
    ND_FORMULA_RECURSIVESPHERE = 47,   /* formulas/recursivesphere_ifs_carve_twist.wgsl */
    ND_FORMULA_TREE = 48,              /* formulas/tree_ifs.wgsl */
    ND_FORMULA_KLEINIANLIMIT = 49,     /* formulas/kleinianlimit_kleinian.wgsl */
    ND_FORMULA_FBMSDF = 50,            /* formulas/fbmsdf_primitive_carve.wgsl */
    ND_FORMULA_SEASCAPE = 51,          /* formulas/seascape_primitive.wgsl */
    ND_FORMULA_DRAGON = 52,            /* formulas/dragon_ifs_twist.wgsl */
    ND_FORMULA_ROMANESCO = 53,         /* formulas/romanesco_ifs.wgsl */
    ND_FORMULA_HEIGHTMAP = 54          /* formulas/heightmap_primitive.wgsl */
} nd_builtin_formula;
#define ND_FORMULA_COUNT 55

#define ND_MAX_CODES 16

/* Normal estimators. */
typedef enum nd_normal_mode {
    ND_NORMAL_ANALYTIC = 0,  /* forward-mode derivative of the DE: the h->0 limit of the central stencil */
    ND_NORMAL_TETRA = 1,     /* reference tetrahedral 4-tap, e = eps(t) (template.wgsl:1781-1804) */
    ND_NORMAL_CENTRAL = 2    /* 6-tap central difference, h = 0.5 * eps(t) */
} nd_normal_mode;

/* Kernel choice for the geometry passes. */
typedef enum nd_kernel {
    ND_KERNEL_AUTO = 0,        /* persistent batched */
    ND_KERNEL_PLAIN = 1,       /* one launch per frame, one lane per pixel, 8x8 tile per wave64 */
    ND_KERNEL_PERSISTENT = 2   /* persistent kernels pulling rays from device queues */
} nd_kernel;

typedef struct nd_march_opts {
    uint32_t normal_mode;       /* nd_normal_mode */
    uint32_t shadow_max_steps;  /* 0: the reference rule min(max_steps, 40) (template.wgsl:2353, 2370) */
    uint32_t shadows;           /* non-zero: march the light-0 soft shadow of every hit */
    uint32_t kernel;            /* nd_kernel */
} nd_march_opts;

/* Pixel rectangle inside the render target. The target is resolution[0] x resolution[1]
 * of the job's Uniforms; w == 0 or h == 0 means the whole target. */
typedef struct nd_rect {
    uint32_t x, y, w, h;
} nd_rect;

/* G-buffer outputs, row-major over the rect (pixel (x, y) of the rect at y * w + x).
 * Any pointer may be NULL (that output is skipped). With device_ptrs == 0 they are host
 * pointers; otherwise device pointers on the context's device.
 *   depth   t at the hit, 0 on a miss (FragOut.depth, template.wgsl:3762-3774)
 *   normal  3 floats per pixel, unit length at hits, 0 on a miss
 *   position  3 floats per pixel: the hit point o + t d as the kernel computed it, 0 on a miss
 *   trap    orbit-trap value at the hit, blended across instances by the combine weight
 *           like the material (template.wgsl:2128-2155); 0 on a miss
 *   hit     1 at hits, 0 at misses (FragOut.normal.w)
 *   shadow  light-0 visibility in [0, 1]; 1 on a miss or when shadows are off
 *   evals   DE evaluations spent on the pixel (march + normal + trap + shadow)
 *   color   3 floats per pixel: fs_simple's colour attachment (FragOut.color.rgb,
 *           template.wgsl:4492-4538): with PART_GEOMETRY_ONLY set in debug_parts the
 *           simple_geometry view (a key light from the camera, ambient and depth cue), else
 *           render_lite (shade_lite's lights, AO and Fresnel at the hit, the sky on a miss);
 *           both with the reference's tetrahedral normal at the hit
 *   albedo  3 floats per pixel: fs_aov's albedo (aov_albedo, template.wgsl:4853-4896):
 *           the hit material's colour mixed toward white by max(reflectiveness,
 *           transparency), the sky clamped to [0, 1] on a miss
 * The colour and albedo outputs take one more pass after the geometry passes; the hotspot
 * and overstep debug views (PART_HOTSPOTS, PART_OVERSTEP) return ND_E_UNSUPPORTED. */
typedef struct nd_gbuffer_out {
    float* depth;
    float* normal;
    float* position;
    float* trap;
    uint8_t* hit;
    float* shadow;
    uint32_t* evals;
    float* color;
    float* albedo;
    uint32_t device_ptrs;
    uint32_t _pad;
} nd_gbuffer_out;

/* One render job of a batch: its own Uniforms (camera, formula, resolution) and outputs. */
typedef struct nd_job {
    const nd_uniforms* uniforms;  /* host pointer, 12,016 bytes */
    nd_rect rect;
    nd_gbuffer_out out;
} nd_job;

typedef struct nd_batch_stats {
    uint64_t rays;          /* primary rays completed */
    uint64_t hits;
    uint64_t de_evals;      /* all scene_de evaluations, counted on the device */
    uint64_t de_evals_march;
    uint64_t de_evals_surface;
    uint64_t de_evals_shadow;
    float ms_total;         /* device time of the whole call (hipEvents) */
    float ms_march;         /* with graph = 1: the march, surface and shadow passes together */
    float ms_surface;       /* with graph = 1: 0 (the replay is timed as a whole) */
    float ms_shadow;        /* with graph = 1: 0 */
    uint32_t kernels;       /* kernel launches issued */
    uint32_t graph;         /* 1: the persistent passes ran as a replayed hipGraph (a launch record
                               seen the batch before is captured; the first of a shape runs pass
                               by pass) */
    uint64_t march_wave_rounds;  /* persistent march: wave rounds, each offering one DE slot to each
                                    of 64 lanes; de_evals_march / (64 x this) is the lane occupancy */
} nd_batch_stats;

/* ---- context ------------------------------------------------------------------ */
nd_status nd_ctx_create(int hip_device, nd_ctx** out);
void nd_ctx_destroy(nd_ctx* ctx);
const char* nd_last_error(const nd_ctx* ctx);
nd_status nd_sync(nd_ctx* ctx);
nd_status nd_set_cancel(nd_ctx* ctx, const volatile int* flag);
/* The HIP stream (hipStream_t) the context launches on, as an opaque pointer, so callers
 * can time or order work against it. */
void* nd_ctx_stream(nd_ctx* ctx);

/* code_to_builtin[c] is the nd_builtin_formula for FractalInstance.code == c (c < n <= 16).
 * The default binding is identity: code c -> builtin c for c < 4. */
nd_status nd_bind_formulas(nd_ctx* ctx, const int32_t* code_to_builtin, uint32_t n);

/* ---- height-map images (the @image formula, formulas/heightmap_primitive.wgsl) ----- */
/* Replaces the reference's HeightMaps (src/gpu/height_maps.zig: sync / decode / upload) and
 * the image_data binding its shader reads (template.wgsl:1110-1193). The host decodes the
 * image file as the reference does with stb_image and passes the pixels; the library does
 * everything after the pixel read. A slot is instance * 4 for an instance's own formula and
 * instance * 4 + 1 + j for its mix-in j (IMAGE_SLOTS_PER_FRACTAL, template.wgsl:1341-1356). */
#define ND_IMAGE_SLOTS 16
#define ND_IMAGE_MAX_EDGE 2048u   /* larger images are box-downsampled by an integer factor */
typedef enum nd_image_format {
    ND_IMAGE_RGBA16 = 0,   /* 4 x uint16 per pixel, straight alpha (stbi_load_16 with 4 channels) */
    ND_IMAGE_RGBA32F = 1   /* 4 x float per pixel (stbi_loadf with 4 channels, an HDR file): RGB is
                              normalised by its finite minimum and maximum, alpha is taken as 1 */
} nd_image_format;
typedef struct nd_image_desc {
    uint32_t width, height;   /* after downsampling */
    uint32_t words;           /* packed words: 4 per texel */
    float slope[4];           /* per channel: the largest change between neighbouring texels, x 1.001 */
} nd_image_desc;
/* Binds the image to a slot (pixels null or w / h 0 unbinds it). A heightmap formula in a
 * slot without an image renders its flat base, as the reference does with no image loaded.
 * ND_E_INVALID for a slot out of range or an empty or oversized (> 2^28 pixels) image. */
nd_status nd_bind_image(nd_ctx* ctx, uint32_t slot, nd_image_format fmt, const void* pixels, uint32_t w,
                        uint32_t h);
/* The texel words a bound image becomes (host only, no context): premultiplied RGBA16 in two
 * words, then the slopes dilated over 8 + 2 and 64 + 2 texels as 8-bit codes per channel
 * (height_maps.zig: decode). words null: only desc is filled (a size query). */
nd_status nd_image_pack(nd_image_format fmt, const void* pixels, uint32_t w, uint32_t h, uint32_t* words,
                        size_t words_cap, nd_image_desc* desc);

/* ---- geometry passes ----------------------------------------------------------- */
/* One frame (or the rect of it) of depth / normal / trap / hit / shadow. */
nd_status nd_render_gbuffer(nd_ctx* ctx, const void* uniforms, size_t uniforms_size,
                            const nd_march_opts* opts, nd_rect rect, const nd_gbuffer_out* out,
                            nd_batch_stats* stats);

/* No counterpart in the reference. nd_render_gbuffer for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * It allocates nothing (no hipMalloc/hipFree, no allocating library routine): the job
 * record, scene, pass counters and hit list live in one library-owned device array per GPU
 * (about 35 MB, held from the library's load on that GPU until the process ends), whose device
 * address nd_ctx_create looks up. Limits: out->device_ptrs must be 1 (ND_E_INVALID otherwise);
 * at most ND_NOALLOC_MAX_RAYS ray slots, 64 per 8x8 tile of the rect, so 512x1024 pixels
 * (ND_E_INVALID beyond); no empty-space accelerator (ND_E_UNSUPPORTED: its grids are
 * allocated); persistent passes of at most ND_NOALLOC_MAX_GRID workgroups; no per-job latency.
 * Results equal nd_render_gbuffer's. Every call on one GPU, from any context, shares that
 * GPU's array and leaves the last call's contents in it: calls that overlap on one GPU race. */
#define ND_NOALLOC_MAX_RAYS (1u << 19)
#define ND_NOALLOC_MAX_GRID 4096u
nd_status nd_render_gbuffer_noalloc(nd_ctx* ctx, const void* uniforms, size_t uniforms_size,
                                    const nd_march_opts* opts, nd_rect rect, const nd_gbuffer_out* out,
                                    nd_batch_stats* stats);

/* Many jobs (frames of different cameras, formulas and resolutions). With
 * ND_KERNEL_PERSISTENT every pass of the whole batch is one launch; with ND_KERNEL_PLAIN
 * each job is its own launch. stats may be NULL. */
nd_status nd_render_batch(nd_ctx* ctx, const nd_job* jobs, uint32_t n_jobs,
                          const nd_march_opts* opts, nd_batch_stats* stats);

/* Per-job device latency of the last nd_render_batch call in milliseconds (the time from
 * the batch start until the job's last ray retired), for p50/p99 reporting. */
nd_status nd_batch_job_latency(nd_ctx* ctx, float* ms_out, uint32_t n_jobs);

/* ---- autotuner ----------------------------------------------------------------- */
#define ND_TUNE_TRACK_LATENCY 1u  /* record per-job completion times (nd_batch_job_latency) */
#define ND_TUNE_NO_GRAPH 2u       /* nd_path_trace: launch each pass of a chunk one by one instead of
                                     replaying the captured hipGraph of the chunk's passes */
#define ND_TUNE_GBUFFER_GRAPH 4u  /* persistent G-buffer passes: capture them into a hipGraph when a batch
                                     repeats the last launch record (same grids, claims, flags and buffers)
                                     and replay it while it stays the same (off by default) */

typedef struct nd_tuning {
    uint32_t waves_per_cu;     /* persistent grid = CUs x waves_per_cu wave64 workgroups for every pass;
                                  0 = each kernel's resident waves per CU, fewer (at least 4) when the
                                  batch has under two claims per wave */
    uint32_t refill_min_idle;  /* a wave pulls new rays when at least this many lanes are idle (default 8) */
    uint32_t chunk_rays;       /* rays a wave claims from the march queue per atomic; shadow claims are
                                  4x; 0 (default) = 96 when every job has the same scene, else 256 */
    uint32_t flags;            /* ND_TUNE_* */
    uint32_t round_cap;        /* 0 = none; otherwise a persistent geometry pass stops each wave after this
                                  many rounds and the call returns ND_E_TIMEOUT (every kernel still exits
                                  normally; outputs of unfinished rays are undefined) */
    uint32_t tile_width;       /* pixels across the tile a wave64 covers: 4, 8, 16 or 32 (tile height
                                  64 / width); 0 = 8 (8x8). Results do not depend on it */
    uint32_t refill_period;    /* the step budget of the persistent march and shadow passes: a wave takes
                                  this many march steps between refills of its idle lanes (a wave whose
                                  lanes are all idle refills at once); 0 = 1 */
    uint32_t mc_pool_log2;     /* the path trace's path pool: a chunk of samples holds up to 2^k paths
                                  (pixels x samples), k 16..27; 0 = the largest k in 24..27 whose pool
                                  fits in a quarter of the device's memory (27 on a 288 GB device). A
                                  larger pool means fewer chunks, each with one tail of bounce passes,
                                  for about 200 B + 68 B per light of memory a path. Adaptive sampling
                                  caps a chunk at half its minimum sample count. Results do not depend
                                  on it beyond float summation order */
} nd_tuning;

nd_status nd_get_tuning(nd_ctx* ctx, nd_tuning* out);
nd_status nd_set_tuning(nd_ctx* ctx, const nd_tuning* t);
/* Times candidate tunings on the given batch and keeps the fastest: coordinate descent over
 * waves per CU, refill threshold, claim size, tile width and refill period, from the current
 * tuning, two sweeps, the best of two runs per candidate. Outputs do not depend on the tuning. */
nd_status nd_autotune(nd_ctx* ctx, const nd_job* jobs, uint32_t n_jobs,
                      const nd_march_opts* opts, nd_tuning* best);

/* ---- adjoint ------------------------------------------------------------------- */
/* Differentiable parameters. */
typedef enum nd_param_id {
    ND_PARAM_BULB_POWER = 0,     /* Mandelbulb p[0] of the first Mandelbulb instance */
    ND_PARAM_BOX_SCALE = 1,      /* Mandelbox p[0] of the first Mandelbox instance */
    ND_PARAM_BOX_MIN_RADIUS = 2, /* Mandelbox p[1] */
    ND_PARAM_CAM_X = 3,          /* camera_pos translation */
    ND_PARAM_CAM_Y = 4,
    ND_PARAM_CAM_Z = 5,
    ND_PARAM_COUNT = 6
} nd_param_id;

/* Gradient of the depth loss L = sum over pixels hit in both (depth - target)^2 with
 * respect to the parameters, through the implicit-function theorem at each hit:
 * dt/dp = -(dDE/dp) / (grad DE . d - epsilon_coefficient).
 * target_depth: per-view target depth (w*h floats each, 0 = miss), device or host as
 * device_ptrs says. depth_out / dtdp_out (ND_PARAM_COUNT floats per pixel) may be NULL.
 * grad_out: ND_PARAM_COUNT doubles (host), loss_out: one double (host). */
typedef struct nd_adjoint_io {
    const float* target_depth;
    float* depth_out;
    float* dtdp_out;
    double* grad_out;
    double* loss_out;
    uint64_t* pixels_used;
    uint32_t device_ptrs;
    uint32_t _pad;
} nd_adjoint_io;

/* A block grid for the mesher's distance pass (MeshParams, src/gpu/mesher.zig:16-27). Point
 * (block * 8 + local + local_offset) of the grid lies at origin + index * cell; local runs
 * over span^3 points per block, x fastest. */
typedef struct nd_mesh_grid {
    float origin[3];
    float cell;
    float iso;              /* subtracted from every DE value */
    uint32_t span;          /* points per block edge, 1..8 (8: fine pass, 1: coarse pass) */
    uint32_t local_offset;  /* added to every local index (block / 2 in the coarse pass) */
    uint32_t device_ptrs;   /* 1: blocks and values are HIP device pointers */
} nd_mesh_grid;

/* cs_mesh_points (mesh.wgsl:49-61) as MeshGpu.run(.points) dispatches it
 * (export_mesh.zig:134-160): values[b * span^3 + l] = scene_de(point l of block b) - iso for
 * n_blocks blocks given as 4 u32 each (x, y, z, unused). kernel_ms (nullable) receives the
 * kernel's time from HIP events on the context's stream. */
nd_status nd_mesh_points(nd_ctx* ctx, const void* uniforms, size_t uniforms_size, const nd_mesh_grid* grid,
                         const uint32_t* blocks, uint32_t n_blocks, float* values, float* kernel_ms);

/* cs_mesh_probe (mesh.wgsl:63-92) as MeshGpu.run(.probe) dispatches it (play_mode.zig:389-402):
 * per point (x, y, z, h) the DE at the point and the tetrahedral gradient with taps |h|,
 * divided by 4 |h|: out = (d, gx, gy, gz). h < 0 evaluates the scene without its carves. */
nd_status nd_mesh_probe(nd_ctx* ctx, const void* uniforms, size_t uniforms_size, const float* points, uint32_t n,
                        float* out, uint32_t device_ptrs);

/* The empty-space accelerator's skip distance (accel_skip, template.wgsl:426-447) at n
 * points (xyz triples, host memory) with the device grid the context builds or reuses for
 * the scene and cascades of `uniforms` (accel_enabled must be set). skip_out: n floats, host
 * memory. No counterpart in the reference: it lets a host or a test check the device grid
 * and its lookup against another implementation. */
nd_status nd_accel_probe(nd_ctx* ctx, const void* uniforms, size_t uniforms_size, const float* pts, uint32_t n,
                         float* skip_out);

nd_status nd_adjoint(nd_ctx* ctx, const nd_uniforms* views, uint32_t n_views,
                     const nd_adjoint_io* io, nd_batch_stats* stats);

/* No counterpart in the reference (it renders, it does not differentiate). The gradient of a
 * colour loss: each view's direct-lighting image - nd_path_trace's colour at the primary hit
 * with Monte Carlo off and the reflection continuation cut (mc_enabled = 0, PART_REFLECTIONS
 * set): ambient 0.15 x AO and every light (point, global, ray) with its soft shadow, Schlick's
 * local share, the hero wavelength's spectrum to RGB and gamut_fit - and its derivative with
 * respect to the same six parameters as nd_adjoint: the hit depth's implicit-function tangent,
 * then dual numbers through the normal, AO, lights and shadows. loss = sum over hit pixels whose
 * target red is >= 0 of |rgb - target|^2. Outputs, all optional: rgb_out (3 floats per pixel,
 * 0 at misses), drgb_out (18 per pixel: [c * 6 + k] = d rgb_c / d theta_k). The scene must keep
 * its material constant over the surface (every visible instance one colour stop or none, the
 * same for all), roughness <= 0.01, no transparency, thin film or subsurface unless their parts
 * are off, no accelerator: ND_E_UNSUPPORTED otherwise;
 * mc_enabled != 0 or reflections on: ND_E_INVALID. With the primary hit cache on the colour is
 * the unquantised normal's. */
typedef struct nd_radiance_io {
    const float* target_rgb;
    float* rgb_out;
    float* drgb_out;
    double* grad_out;
    double* loss_out;
    uint64_t* pixels_used;
    uint32_t device_ptrs;
    uint32_t _pad;
} nd_radiance_io;

nd_status nd_radiance_adjoint(nd_ctx* ctx, const nd_uniforms* views, uint32_t n_views,
                              const nd_radiance_io* io, nd_batch_stats* stats);

/* ---- wavefront Monte Carlo path trace -------------------------------------------- */
typedef struct nd_image_out {
    float* color;      /* RGB float per pixel: the average over spp samples */
    float* depth;      /* primary hit depth, 0 on a miss */
    float* normal;     /* 3 floats per pixel */
    uint32_t device_ptrs;
    uint32_t _pad;
} nd_image_out;

typedef struct nd_trace_stats {
    uint64_t paths;          /* camera paths traced (pixels x spp) */
    uint64_t segments;       /* extend rays over all bounces */
    uint64_t shadow_rays;
    uint64_t de_evals;
    float ms_total;
    uint32_t max_depth_reached;  /* the depth limit clamp(max_reflection_bounces + 8, 1, 16) */
    uint32_t _pad;
} nd_trace_stats;

/* Path-traces rect with spp samples per pixel: sample i uses mc_sample = u.mc_sample + i,
 * the host's per-sample stamp (fr.zig:3761-3775), and trace_path's transport
 * (template.wgsl:3819-4137): opaque, transmissive (refraction with the inside march,
 * Cauchy dispersion), thin-film and subsurface materials, the Monte Carlo light pick, the
 * primary hit cache (primary_cache_key) and adaptive sampling (adapt_enabled). out->color
 * (required) receives the mean over the samples after gamut_fit; depth and normal come from
 * sample 0 (tetrahedral normal). Fog, photons, depth of field, the approximations and ray
 * lights return ND_E_UNSUPPORTED, so the host can fall back to its WGSL pass. The cancel
 * flag is checked between chunks of samples (ND_E_CANCELLED). */
nd_status nd_path_trace(nd_ctx* ctx, const void* uniforms, size_t uniforms_size, uint32_t spp,
                        nd_rect rect, const nd_image_out* out, nd_trace_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* ND_HIP_H */
