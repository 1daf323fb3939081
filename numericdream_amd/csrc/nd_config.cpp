// nd_config.cpp - Uniforms construction for the five workloads (host side).
//
// Mirrors buildUniforms / stampEyeBasis (src/app/export_image.zig:90-167) and the field
// defaults of the Zig extern structs (src/gpu/fractal_renderer.zig:85-363).
#include "nd_config.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

// splitmix64 seeds xoshiro256** (Blackman & Vigna); doubles from the top 53 bits.
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) {
            x += 0x9e3779b97f4a7c15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
    int integer(int lo, int hi) { return lo + (int)(uniform() * (double)(hi - lo + 1)); }  // [lo, hi]
};

// Stream of job `index`: independent per job so any one block can be rebuilt alone.
uint64_t job_seed(uint64_t seed, uint32_t config, uint32_t index) {
    return seed * 0x100000001b3ull ^ ((uint64_t)config << 32) ^ (uint64_t)index * 0x9e3779b97f4a7c15ull;
}

void set3(float* d, float a, float b, float c) { d[0] = a; d[1] = b; d[2] = c; }

void color_stop_defaults(nd_color_stop* c) {
    std::memset(c, 0, sizeof(*c));
    c->inner_max_steps = 64;
    c->film_thickness = 400;
    c->film_ior = 1.8f;
    c->film_angle_scale = 1;
    c->film_perturb_scale = 0.2f;
    c->subsurface_radius = 0.1f;
}

// FractalInstance defaults (fractal_renderer.zig:119-145) for a used slot.
void instance_defaults(nd_fractal_instance* in) {
    std::memset(in, 0, sizeof(*in));
    in->step_safety = 1;
    set3(in->scale, 1, 1, 1);
    in->code = ND_NO_CODE;
    in->hybrid_base_iters = 2;
    in->hybrid_total_iters = 12;
    set3(in->trap_box, 1, 1, 1);
    in->trap_radius = 1;
    in->trap_span = 1.5f;
    for (int j = 0; j < ND_MAX_MIXINS; ++j) {
        in->mixins[j].iterations = 1;
        in->mixins[j].code = ND_NO_CODE;
    }
    for (int j = 0; j < ND_MAX_COLOR_STOPS; ++j) color_stop_defaults(&in->colors[j]);
    in->frame[0][0] = in->frame[1][1] = in->frame[2][2] = 1;
}

// Point on the sphere of radius r, uniform in direction: z = U[-1,1], phi = U[0, 2pi).
void sphere_point(Rng& g, double r, float out[3]) {
    double z = g.uniform(-1.0, 1.0);
    double phi = g.uniform(0.0, 2.0 * M_PI);
    double s = std::sqrt(std::fmax(0.0, 1.0 - z * z));
    out[0] = (float)(r * s * std::cos(phi));
    out[1] = (float)(r * z);
    out[2] = (float)(r * s * std::sin(phi));
}

void black_white_strip(nd_fractal_instance* in) {
    in->color_count = 2;
    set3(in->colors[0].color, 0, 0, 0);
    in->colors[0].position = 0;
    set3(in->colors[1].color, 1, 1, 1);
    in->colors[1].position = 1;
}

// Sphere orbit trap of radius 1, ITERATION mode: the trap value is (index of the iteration
// whose z came closest to the unit sphere + 1) / iterations.
void sphere_trap(nd_fractal_instance* in) {
    in->trap_shape = 1;
    in->trap_radius = 1;
    in->trap_mode = 4;
}

void mandelbulb(nd_fractal_instance* in, float power, float iters) {
    in->code = (float)ND_FORMULA_MANDELBULB;
    in->params[0] = power;
    in->params[1] = iters;
}

// amazingbox_box.wgsl @param defaults with the workload's overrides
void mandelbox(nd_fractal_instance* in, float scale, float min_r, float fix_r, float limit, float iters) {
    in->code = (float)ND_FORMULA_MANDELBOX;
    in->params[0] = scale;
    in->params[1] = min_r;
    in->params[2] = fix_r;
    in->params[3] = limit;
    in->params[4] = iters;
    in->params[8] = 1;  // CScale
}

void march_defaults(nd_uniforms* u, float w, float h, float max_steps, float max_dist) {
    u->resolution[0] = w;
    u->resolution[1] = h;
    u->max_steps = max_steps;
    u->max_dist = max_dist;
    u->high_quality = 1;  // export path (export_image.zig:128)
    u->debug_parts = (float)ND_PART_GEOMETRY_ONLY;
}

}  // namespace

extern "C" {

uint32_t nd_config_job_count(uint32_t config) {
    switch (config) {
        case 1: return 1;
        case 2: return 16;
        case 3: return 1;
        case 4: return 256;
        case 5: return 8;
    }
    return 0;
}

void nd_uniforms_defaults(nd_uniforms* u) {
    std::memset(u, 0, sizeof(*u));
    u->max_reflection_bounces = 1;
    u->epsilon_coefficient = 0.0005f;
    u->epsilon_floor = 0.00005f;
    u->hq_footprint_budget_px = 2.0f;
    u->tile_scale[0] = u->tile_scale[1] = 1;
    u->focus_distance = 3.2f;
    u->focus_range = 0.5f;
    u->refine_fast = 4;
    u->refine_hq = 20;
    for (int i = 0; i < ND_MAX_INSTANCES; ++i) {
        std::memset(&u->instances[i], 0, sizeof(nd_fractal_instance));
        u->instances[i].code = ND_NO_CODE;  // unused slot (fr.zig:55, 2888)
    }
    u->fog_samples = 1;
    u->slice_zoom = 2.0f;
    u->accel_safety = 1.0f;
    for (int i = 0; i < ND_MAX_WARPS; ++i) {
        set3(u->warps[i].extent, 1, 1, 1);
        u->warps[i].safety = 1;
        u->warps[i].frame[0][0] = u->warps[i].frame[1][1] = u->warps[i].frame[2][2] = 1;
    }
    u->photon_radius = 0.05f;
    u->photon_cell = 0.1f;
    u->photon_intensity = 1;
    u->sky.intensity = 1;
    u->sky.falloff = 1;
    set3(u->sky.zenith, 0.02f, 0.02f, 0.05f);
    u->sky.sun_cos = -2;
    set3(u->sky.horizon, 0.04f, 0.035f, 0.07f);
    set3(u->sky.ground, 0.06f, 0.05f, 0.09f);
    set3(u->sky.sun_color, 1, 1, 1);
    set3(u->sky.sun_dir, 0, 1, 0);
    u->photon_cap_surface = 8;
    u->photon_fixed_unit = 1;
    u->photon_volume_scale = 1;
    u->fft_box_radius = 2.5f;
    u->fft_cloud_density = 4.0f;
    u->fft_lowpass = 1.0f;
    u->photon_bounce_scale = 4;
    for (int i = 0; i < ND_MAX_PARTICLE_SYSTEMS; ++i) {
        u->particle_systems[i].glow = 1;
        u->particle_systems[i].glow_extent = 3;
    }
    u->approx_caustic_strength = 1;
    u->approx_caustic_scale = 0.3f;
    u->approx_bounce = 0.35f;
    u->approx_tint = 1;
    u->approx_lod_start = 2;
    u->approx_lod_strength = 1.5f;
    u->adapt_threshold = 0.02f;
    u->adapt_min_samples = 16;
    u->flat_bailout = 16;
    u->flat_band_density = 0.25f;
    u->flat_light_angle = 135;
    u->flat_contour_density = 2;
    u->flat_slab_slices = 1;
    u->flat_slab_depth = 0.5f;
    u->flat_slab_opacity = 0.35f;
    u->flat_samples = 1;
    u->flat_halo = 0.35f;
    u->mc_inner_samples = 1;
    for (int i = 0; i < ND_MAX_LIGHTS / 4; ++i)
        for (int k = 0; k < 4; ++k) u->fog_vis_light_layer[i][k] = -1;
}

void nd_camera_look_at(nd_uniforms* u, const float pos[3], const float target[3], float fov_deg) {
    double f[3] = {(double)target[0] - pos[0], (double)target[1] - pos[1], (double)target[2] - pos[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    for (double& x : f) x /= fl;
    double wu[3] = {0, 1, 0};
    if (std::fabs(f[1]) > 0.99) { wu[1] = 0; wu[2] = 1; }
    double r[3] = {f[1] * wu[2] - f[2] * wu[1], f[2] * wu[0] - f[0] * wu[2], f[0] * wu[1] - f[1] * wu[0]};
    double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (double& x : r) x /= rl;
    double up[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
    double ul = std::sqrt(up[0] * up[0] + up[1] * up[1] + up[2] * up[2]);
    for (double& x : up) x /= ul;
    double k = std::tan(0.5 * (double)fov_deg * M_PI / 180.0);
    for (int i = 0; i < 3; ++i) {
        u->camera_pos[i] = pos[i];
        u->camera_forward[i] = (float)f[i];
        u->camera_right[i] = (float)(r[i] * k);
        u->camera_up[i] = (float)(up[i] * k);
    }
}

void nd_accel_setup(nd_uniforms* u, uint32_t res, uint32_t levels, float safety) {
    // placeCascades (accel.zig:34-58) in the same f32 arithmetic
    const int n = std::min(std::max((int)levels, 1), ND_MAX_CASCADES);
    const float res_f = (float)std::max(res, 1u);
    const float coarsest = std::max(2.0f * u->max_dist / res_f, 1e-6f);
    for (int l = 0; l < ND_MAX_CASCADES; ++l)
        for (int k = 0; k < 4; ++k) u->accel_params[l][k] = 0.0f;
    for (int l = 0; l < n; ++l) {
        const float cell = coarsest / (float)(1u << (n - 1 - l));
        const float half = 0.5f * res_f * cell;
        for (int a = 0; a < 3; ++a) u->accel_params[l][a] = std::floor((u->camera_pos[a] - half) / cell) * cell;
        u->accel_params[l][3] = cell;
    }
    u->accel_enabled = 1.0f;
    u->accel_levels = (float)n;
    u->accel_res = (float)res;
    u->accel_safety = safety;
    u->accel_z_offset = 0.0f;
}

void nd_set_pixel_epsilon(nd_uniforms* u, float fov_deg, float fraction) {
    double theta_px = 2.0 * std::tan(0.5 * (double)fov_deg * M_PI / 180.0) / (double)u->resolution[1];
    u->epsilon_coefficient = (float)((double)fraction * theta_px);
    u->epsilon_floor = 0.0f;
}

nd_status nd_config_build(uint32_t config, uint32_t index, uint64_t seed, nd_uniforms* u) {
    if (!u) return ND_E_INVALID;
    if (index >= nd_config_job_count(config)) return ND_E_INVALID;
    if (seed == 0) seed = ND_DEFAULT_SEED;
    nd_uniforms_defaults(u);
    Rng g(job_seed(seed, config, index));
    const float origin[3] = {0, 0, 0};
    const float fov = 60.0f;

    if (config == 1) {
        march_defaults(u, 640, 360, 256, 10);
        u->instance_count = 1;
        nd_fractal_instance* in = &u->instances[0];
        instance_defaults(in);
        mandelbulb(in, 8, 8);
        sphere_trap(in);
        black_white_strip(in);
        const float cam[3] = {0, 0, -2.5f};
        nd_camera_look_at(u, cam, origin, fov);
        nd_set_pixel_epsilon(u, fov, 1e-3f);
        return ND_OK;
    }
    if (config == 2) {
        march_defaults(u, 1920, 1080, 256, 10);
        u->instance_count = 1;
        nd_fractal_instance* in = &u->instances[0];
        instance_defaults(in);
        mandelbox(in, -1.5f, 0.5f, 1.0f, 1.0f, 15);
        sphere_trap(in);
        black_white_strip(in);
        float cam[3];
        sphere_point(g, 4.0, cam);
        nd_camera_look_at(u, cam, origin, fov);
        nd_set_pixel_epsilon(u, fov, 1e-3f);
        u->light_count = 1;
        nd_light* L = &u->lights[0];
        set3(L->color, 1, 1, 1);
        L->brightness = 1;
        set3(L->position_or_direction, 0.577f, 0.577f, -0.577f);
        L->light_type = 1;  // LIGHT_GLOBAL
        L->shadow_softness = 16;
        L->cast_shadows = 1;
        set3(L->ray_direction, 0, -1, 0);
        L->waist = 0.05f;
        return ND_OK;
    }
    if (config == 3) {
        march_defaults(u, 1920, 1080, 256, 24);
        u->debug_parts = 0;
        u->mc_enabled = 1;
        u->instance_count = 4;
        for (int i = 0; i < 4; ++i) {
            nd_fractal_instance* in = &u->instances[i];
            instance_defaults(in);
            if (i > 0) {
                in->combine_mode = 1;  // COMBINE_SMOOTH_MIN
                in->blend_k = 0.1f;
            }
            sphere_trap(in);
            in->trap_mode = 0;  // min distance
            in->color_count = 16;
            for (int s = 0; s < 16; ++s) {
                nd_color_stop* c = &in->colors[s];
                c->position = (float)s / 15.0f;
                for (int k = 0; k < 3; ++k) c->color[k] = (float)g.uniform(0.2, 0.9);
                c->roughness = (float)g.uniform(0.05, 1.0);
                c->reflectiveness = g.uniform() < 0.5 ? 0.0f : 1.0f;  // "metallic" 0/1
                c->glossiness = 0.3f;
                c->ior = 1.5f;
            }
        }
        mandelbulb(&u->instances[0], 8, 6);
        mandelbox(&u->instances[1], 2.0f, 0.5f, 1.0f, 1.0f, 12);
        nd_fractal_instance* m = &u->instances[2];
        m->code = (float)ND_FORMULA_MENGER3;
        m->params[0] = 5;
        m->params[1] = m->params[2] = m->params[3] = 1;
        m->params[4] = 3;
        nd_fractal_instance* k = &u->instances[3];
        k->code = (float)ND_FORMULA_KLEINIAN;
        k->params[0] = 1;                                   // FoldingLimit
        k->params[1] = k->params[2] = k->params[3] = 1;     // Center (1,1,1): "box size (1,1,1)"
        k->params[4] = 1;                                   // Radius
        k->params[5] = 2;                                   // CellSize
        k->params[6] = 12;                                  // Iterations
        k->params[7] = 1;                                   // CScale
        const float cam[3] = {0.0f, 1.5f, -9.0f};
        nd_camera_look_at(u, cam, origin, fov);
        nd_set_pixel_epsilon(u, fov, 1e-3f);
        u->light_count = 1;
        nd_light* L = &u->lights[0];
        set3(L->color, 1, 1, 1);
        L->brightness = 20;
        set3(L->position_or_direction, 2, 3, -2);
        L->light_type = 0;  // LIGHT_POINT
        L->shadow_softness = 16;
        L->cast_shadows = 1;
        set3(L->ray_direction, 0, -1, 0);
        L->waist = 0.05f;
        u->max_reflection_bounces = 4;
        return ND_OK;
    }
    if (config == 4) {
        int side = (int)std::lround(std::exp(g.uniform(std::log(64.0), std::log(1024.0))));
        march_defaults(u, (float)side, (float)side, 256, 16);
        u->instance_count = 1;
        nd_fractal_instance* in = &u->instances[0];
        instance_defaults(in);
        bool bulb = g.uniform() < 0.5;
        float iters = (float)g.integer(8, 20);
        if (bulb) {
            mandelbulb(in, (float)g.uniform(2.0, 12.0), iters);
        } else {
            // uniform over [-2.5,-1.2] u [1.5,3] (lengths 1.3 and 1.5)
            double x = g.uniform(0.0, 2.8);
            float scale = (float)(x < 1.3 ? -2.5 + x : 1.5 + (x - 1.3));
            mandelbox(in, scale, 0.5f, 1.0f, 1.0f, iters);
        }
        sphere_trap(in);
        black_white_strip(in);
        float cam[3];
        sphere_point(g, g.uniform(2.5, 5.0), cam);
        nd_camera_look_at(u, cam, origin, fov);
        nd_set_pixel_epsilon(u, fov, 1e-3f);
        return ND_OK;
    }
    if (config == 5) {
        march_defaults(u, 512, 512, 256, 10);
        u->instance_count = 2;
        nd_fractal_instance* a = &u->instances[0];
        instance_defaults(a);
        mandelbulb(a, 8, 8);
        set3(a->offset, -0.75f, 0, 0);
        set3(a->scale, 0.7f, 0.7f, 0.7f);
        sphere_trap(a);
        nd_fractal_instance* b = &u->instances[1];
        instance_defaults(b);
        mandelbox(b, -1.5f, 0.5f, 1.0f, 1.0f, 12);
        set3(b->offset, 0.75f, 0, 0);
        set3(b->scale, 0.3f, 0.3f, 0.3f);
        b->combine_mode = 0;  // union
        sphere_trap(b);
        float cam[3];
        sphere_point(g, 3.0, cam);
        nd_camera_look_at(u, cam, origin, fov);
        nd_set_pixel_epsilon(u, fov, 1.0f);  // one pixel footprint (see nd_config.h)
        return ND_OK;
    }
    return ND_E_INVALID;
}

}  // extern "C"
