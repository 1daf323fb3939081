// nd_mc.hip - wavefront Monte Carlo path trace for gfx950.
//
// Transport of the reference's trace_path (template.wgsl:3819-4137): 8 hero-stratified
// wavelengths, the procedural sky, point / global lights with soft shadows tinted by
// transparent occluders, the Monte Carlo light pick, Blinn-Phong shading, Schlick reflection
// and refraction off roughness-perturbed normals with the inside march through solids,
// Cauchy dispersion on a picked wavelength, one cosine-sampled indirect bounce at the
// primary hit, Russian roulette from depth 4.
//
// Wavefront schedule (the reference's cs_wf_* sequence, template.wgsl:4898-5407, as a
// device-side loop): per bounce
//   extend    persistent march of every live path (rays refilled per lane)
//   shade     normal, material, scatter; the light-0..n shadow rays and the primary's
//             indirect ray are queued with their spectral weights, the continuation is
//             compacted into the next live queue (one atomic per wave and queue)
//   shadow    persistent visibility march of the shadow queue
//   indirect  (bounce 0) persistent march of the indirect queue, then its shading
// then resolve: per pixel, the average over samples of gamut_fit(spectrum -> RGB).
// Path state is SoA in HBM. Random numbers hash the FP32 bits of the same seeds as the
// reference (hash12w), computed without fused multiply-adds.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "nd_accel.hpp"
#include "nd_ctx.h"
#include "nd_de.hpp"
#include "nd_internal.h"
#include "nd_march_state.hpp"
#include "nd_rays.hpp"

namespace nd {

// ----------------------------------------------------------------------------- scene
struct McStop {
    float color[3], position, glossiness, transparency, reflectiveness, ior, roughness, abbe, inner_max_steps;
    float film_thickness, film_ior, film_strength, film_angle_scale, film_perturb, film_perturb_scale;
    float subsurface, subsurface_radius;
};
struct McInst {
    int color_count, trap_repeat;
    float trap_span, trap_offset;
    McStop stops[16];
};
struct McLight {
    int type;  // 0 point, 1 global, 2 ray (a beam from v along axis)
    int cast;
    float color[3];  // colour * brightness
    float v[3];      // position or normalised direction
    float k;         // penumbra
    float axis[3];   // ray light: the beam's unit axis
    float waist;     // ray light: beam radius at the source
    float tan_spread;  // ray light: tan of the half-angle of its widening
};
struct McScene {
    McInst inst[ND_DEV_MAX_INST];
    McLight lights[8];
    int light_count;
    int shadows_on, specular_on, ao_on, reflections_on, indirect_on, sky_on;
    int strips_on, refraction_on, dispersion_on, iridescence_on, subsurface_on;  // apply_part_mask
    int sss;           // some stop has subsurface: the shade pass runs its SSS variant
    int sss_steps;     // min(max_steps, SSS_PATH_MAX_STEPS = 128)
    int transp;        // scene_has_transparency: shadow rays cross transparent occluders
    int transmissive;  // some visible stop has transparency: paths can enter solids
    int light_pick;    // mc_enabled and mc_light_pick
    int cache_on;      // primary_cache_key: the primary hit cache
    int adapt_on;      // adapt_enabled and mc_enabled
    float adapt_row_width, adapt_min_samples, adapt_threshold;
    float disp_soft;   // radians(0.5 * photon_dispersion_soft)
    int max_depth;
    int shadow_steps;
    float sky_zenith[3], sky_horizon[3], sky_ground[3], sky_sun_color[3], sky_sun_dir[3];
    float sky_yaw_c, sky_yaw_s, sky_inv_falloff, sky_sun_cos, sky_sun_intensity, sky_intensity;
    int sky_sun;
    float max_dist;
    int simple_view;   // PART_GEOMETRY_ONLY set: fs_simple draws simple_geometry, else render_lite
    // depth of field (cs_wf_generate, template.wgsl:5150-5182): a pixel whose centre ray's hit
    // lies outside the focus range takes dof_n lens samples per sample, averaged by the resolve
    int dof_on;        // dof_enabled, aperture > DOF_MIN_APERTURE, PART_DEPTH_OF_FIELD on
    int slots;         // path slots per pixel sample: the wavefront's inner count with DOF, else 1
    int dof_n;         // lens samples of an out-of-focus pixel: min(mc_inner_count(32), slots)
    float aperture, focus_distance, focus_range;
};

// ----------------------------------------------------------------------------- random numbers
NDD uint32_t hash_u32(float px, float py) {
    uint32_t x = __float_as_uint(px), y = __float_as_uint(py);
    const uint32_t M = 1664525u, A = 1013904223u;
    x = x * M + A;
    y = y * M + A;
    x += y * M;
    y += x * M;
    x ^= x >> 16;
    y ^= y >> 16;
    x += y * M;
    y += x * M;
    x ^= x >> 16;
    y ^= y >> 16;
    return x ^ y;
}
NDD float hash12w(float px, float py) { return (float)hash_u32(px, py) * 2.3283064365386963e-10f; }
NDD float ffract(float x) { return x - floorf(x); }
// interleaved gradient noise (hash12, template.wgsl:2413-2415):
// fract(52.9829189 fract(dot(p, (0.06711056, 0.00583715)))), one rounding per operation
NDD float hash12(float px, float py) {
    return ffract(mul_rn(52.9829189f, ffract(add_rn(mul_rn(px, 0.06711056f), mul_rn(py, 0.00583715f)))));
}
// seed arithmetic without contraction, as the shader's separate multiply and add
NDD float mad_rn(float a, float b, float c) { return add_rn(mul_rn(a, b), c); }

// ----------------------------------------------------------------------------- spectra
struct Sp8 { float v[8]; };
NDD Sp8 sp_splat(float a) { Sp8 s; for (int k = 0; k < 8; ++k) s.v[k] = a; return s; }
NDD float sp_max(const Sp8& a) { float m = a.v[0]; for (int k = 1; k < 8; ++k) m = fmaxf(m, a.v[k]); return m; }

// RGB -> spectrum basis of a path's 8 wavelengths (spectral_setup, template.wgsl:816-850)
struct Basis { float r[8], g[8], b[8]; };
NDD void lambdas_of(float hero, float* l) {
    const float base = hero - floorf(hero);
    for (int k = 0; k < 8; ++k) {
        float f = base + 0.125f * (float)k;
        f = f - floorf(f);
        l[k] = 380.0f + 350.0f * f;
    }
}
NDD Basis basis_of(float hero) {
    float l[8];
    lambdas_of(hero, l);
    Basis B;
    for (int k = 0; k < 8; ++k) {
        float blue = 1.0f / (1.0f + expf((l[k] - 498.09f) / 13.50f));
        float red = 1.0f / (1.0f + expf((600.31f - l[k]) / 8.00f));
        float t = (l[k] - 530.0f) / 25.0f;
        float green = 2.49f * expf(-0.5f * t * t);
        float is = 1.0f / (red + green + blue);
        float a = red * is, b = green * is, c = blue * is;
        B.r[k] = a * 0.99575300f + b * -0.00755455f + c * 0.02732810f;
        B.g[k] = a * -0.00656071f + b * 1.00830220f + c * -0.00674214f;
        B.b[k] = a * 0.01080772f + b * -0.00074766f + c * 0.97941404f;
    }
    return B;
}
NDD Sp8 from_rgb(const Basis& B, float r, float g, float b) {
    Sp8 s;
    for (int k = 0; k < 8; ++k) s.v[k] = fmaxf(r * B.r[k] + g * B.g[k] + b * B.b[k], 0.f);
    return s;
}
NDD float cie_lobe(float l, float mu, float slo, float shi) {
    float s = l < mu ? slo : shi;
    float u = (l - mu) / s;
    return expf(-0.5f * u * u);
}
// spectrum -> RGB with the CIE-fit weights, then gamut_fit (template.wgsl:880-908)
NDD void to_rgb_fit(float hero, const Sp8& s, float* rgb) {
    float l[8];
    lambdas_of(hero, l);
    float wr[8], wg[8], wb[8], sr = 0.f, sg = 0.f, sb = 0.f;
    for (int k = 0; k < 8; ++k) {
        float x = 1.056f * cie_lobe(l[k], 599.8f, 37.9f, 31.0f) + 0.362f * cie_lobe(l[k], 442.0f, 16.0f, 26.7f) -
                  0.065f * cie_lobe(l[k], 501.1f, 20.4f, 26.2f);
        float y = 0.821f * cie_lobe(l[k], 568.8f, 46.9f, 40.5f) + 0.286f * cie_lobe(l[k], 530.9f, 16.3f, 31.1f);
        float z = 1.217f * cie_lobe(l[k], 437.0f, 11.8f, 36.0f) + 0.681f * cie_lobe(l[k], 459.0f, 26.0f, 13.8f);
        wr[k] = 3.2404542f * x - 1.5371385f * y - 0.4985314f * z;
        wg[k] = -0.9692660f * x + 1.8760108f * y + 0.0415560f * z;
        wb[k] = 0.0556434f * x - 0.2040259f * y + 1.0572252f * z;
        sr += wr[k]; sg += wg[k]; sb += wb[k];
    }
    float c0 = 0.f, c1 = 0.f, c2 = 0.f;
    for (int k = 0; k < 8; ++k) { c0 += s.v[k] * wr[k]; c1 += s.v[k] * wg[k]; c2 += s.v[k] * wb[k]; }
    c0 /= sr; c1 /= sg; c2 /= sb;
    const float y = c0 * 0.2126f + c1 * 0.7152f + c2 * 0.0722f;
    if (y <= 0.f) { rgb[0] = rgb[1] = rgb[2] = 0.f; return; }
    const float m = fminf(c0, fminf(c1, c2));
    if (m < 0.f) {
        const float f = y / (y - m);
        c0 = y + (c0 - y) * f; c1 = y + (c1 - y) * f; c2 = y + (c2 - y) * f;
    }
    rgb[0] = c0; rgb[1] = c1; rgb[2] = c2;
}

// ----------------------------------------------------------------------------- sky & materials
NDD V3<float> sky_color(const McScene& S, const V3<float>& dir) {
    if (!S.sky_on) return mk3(0.f, 0.f, 0.f);
    V3<float> d = mk3(S.sky_yaw_c * dir.x - S.sky_yaw_s * dir.z, dir.y, S.sky_yaw_s * dir.x + S.sky_yaw_c * dir.z);
    V3<float> c;
    if (d.y >= 0.f) {
        float t = powf(fclamp(d.y, 0.f, 1.f), S.sky_inv_falloff);
        c = mk3(S.sky_horizon[0] + (S.sky_zenith[0] - S.sky_horizon[0]) * t,
                S.sky_horizon[1] + (S.sky_zenith[1] - S.sky_horizon[1]) * t,
                S.sky_horizon[2] + (S.sky_zenith[2] - S.sky_horizon[2]) * t);
    } else {
        float t = fclamp(-d.y, 0.f, 1.f);
        c = mk3(S.sky_horizon[0] + (S.sky_ground[0] - S.sky_horizon[0]) * t,
                S.sky_horizon[1] + (S.sky_ground[1] - S.sky_horizon[1]) * t,
                S.sky_horizon[2] + (S.sky_ground[2] - S.sky_horizon[2]) * t);
    }
    if (S.sky_sun) {
        float dd = d.x * S.sky_sun_dir[0] + d.y * S.sky_sun_dir[1] + d.z * S.sky_sun_dir[2];
        float e0 = S.sky_sun_cos, e1 = S.sky_sun_cos + (1.f - S.sky_sun_cos) * 0.15f;
        float x = fclamp((dd - e0) / (e1 - e0), 0.f, 1.f);
        float w = S.sky_sun_intensity * x * x * (3.f - 2.f * x);
        c = mk3(c.x + S.sky_sun_color[0] * w, c.y + S.sky_sun_color[1] * w, c.z + S.sky_sun_color[2] * w);
    }
    return mk3(fmaxf(c.x * S.sky_intensity, 0.f), fmaxf(c.y * S.sky_intensity, 0.f), fmaxf(c.z * S.sky_intensity, 0.f));
}

struct Mat {
    float color[3], glossiness, transparency, reflectiveness, ior, roughness, abbe, inner_max_steps;
    float film_thickness, film_ior, film_strength, film_angle_scale, film_perturb, film_perturb_scale;
    float subsurface, subsurface_radius;
};
// default_material (template.wgsl:1965-1978)
NDD Mat mat_default() {
    Mat m;
    m.color[0] = 0.8f; m.color[1] = 0.8f; m.color[2] = 0.85f;
    m.glossiness = 0.3f; m.transparency = 0.f; m.reflectiveness = 0.f; m.ior = 1.5f; m.roughness = 0.f;
    m.abbe = 0.f; m.inner_max_steps = 64.f;
    m.film_thickness = 400.f; m.film_ior = 1.8f; m.film_strength = 0.f; m.film_angle_scale = 1.f;
    m.film_perturb = 0.f; m.film_perturb_scale = 0.2f;
    m.subsurface = 0.f; m.subsurface_radius = 0.1f;  // DEFAULT_SSS_RADIUS
    return m;
}
NDD Mat mat_stop(const McStop& s) {
    Mat m;
    for (int k = 0; k < 3; ++k) m.color[k] = s.color[k];
    m.glossiness = s.glossiness; m.transparency = s.transparency; m.reflectiveness = s.reflectiveness; m.ior = s.ior;
    m.roughness = s.roughness; m.abbe = s.abbe; m.inner_max_steps = s.inner_max_steps;
    m.film_thickness = s.film_thickness; m.film_ior = s.film_ior; m.film_strength = s.film_strength;
    m.film_angle_scale = s.film_angle_scale; m.film_perturb = s.film_perturb; m.film_perturb_scale = s.film_perturb_scale;
    m.subsurface = s.subsurface; m.subsurface_radius = s.subsurface_radius;
    return m;
}
// apply_part_mask (template.wgsl:394-402) for the parts this path reads
NDD Mat part_mask(Mat m, const McScene& S) {
    if (!S.strips_on) { m.color[0] = 0.72f; m.color[1] = 0.72f; m.color[2] = 0.73f; }  // SIMPLE_ALBEDO
    if (!S.reflections_on) m.reflectiveness = 0.f;
    if (!S.refraction_on) m.transparency = 0.f;
    if (!S.dispersion_on) m.abbe = 0.f;
    if (!S.iridescence_on) m.film_strength = 0.f;
    if (!S.subsurface_on) m.subsurface = 0.f;
    return m;
}
// mix_material(a, b, h) = mix(b, a, h) (template.wgsl:1980-2001)
NDD Mat mat_mix(const Mat& a, const Mat& b, float h) {
    Mat r;
    for (int k = 0; k < 3; ++k) r.color[k] = b.color[k] * (1.f - h) + a.color[k] * h;
    r.glossiness = b.glossiness * (1.f - h) + a.glossiness * h;
    r.reflectiveness = b.reflectiveness * (1.f - h) + a.reflectiveness * h;
    r.ior = b.ior * (1.f - h) + a.ior * h;
    r.roughness = b.roughness * (1.f - h) + a.roughness * h;
    r.transparency = b.transparency * (1.f - h) + a.transparency * h;
    r.abbe = b.abbe * (1.f - h) + a.abbe * h;
    r.inner_max_steps = b.inner_max_steps * (1.f - h) + a.inner_max_steps * h;
    r.film_thickness = b.film_thickness * (1.f - h) + a.film_thickness * h;
    r.film_ior = b.film_ior * (1.f - h) + a.film_ior * h;
    r.film_strength = b.film_strength * (1.f - h) + a.film_strength * h;
    r.film_angle_scale = b.film_angle_scale * (1.f - h) + a.film_angle_scale * h;
    r.film_perturb = b.film_perturb * (1.f - h) + a.film_perturb * h;
    r.film_perturb_scale = b.film_perturb_scale * (1.f - h) + a.film_perturb_scale * h;
    r.subsurface = b.subsurface * (1.f - h) + a.subsurface * h;
    r.subsurface_radius = b.subsurface_radius * (1.f - h) + a.subsurface_radius * h;
    return r;
}
// sample_gradient + orbit_trap_to_strip (template.wgsl:2003-2035)
NDD Mat strip_mat(const McInst& I, float trap) {
    float x = trap / fmaxf(I.trap_span, 1e-4f) + I.trap_offset;
    float t;
    if (I.trap_repeat == 1) t = x - floorf(x);
    else if (I.trap_repeat == 2) { float f = x * 0.5f - floorf(x * 0.5f); t = 1.f - fabsf(f * 2.f - 1.f); }
    else t = fclamp(x, 0.f, 1.f);
    const int count = min(I.color_count, 16);
    if (count <= 0) return mat_default();
    if (count == 1 || t <= I.stops[0].position) return mat_stop(I.stops[0]);
    if (t >= I.stops[count - 1].position) return mat_stop(I.stops[count - 1]);
    for (int i = 0; i < count - 1; ++i) {
        const McStop &a = I.stops[i], &b = I.stops[i + 1];
        if (t >= a.position && t <= b.position) {
            float f = (t - a.position) / fmaxf(b.position - a.position, 1e-4f);
            return mat_mix(mat_stop(a), mat_stop(b), 1.f - f);
        }
    }
    return mat_stop(I.stops[count - 1]);
}
// hit_material (template.wgsl:2128-2155)
template <bool EXT>
NDD Mat hit_material(const DevScene& sc, const McScene& S, const V3<float>& p) {
    bool first = true;
    float d = sc.empty_de;
    Mat m = mat_default();
    for (int i = 0; i < sc.count; ++i) {
        const DevInst& in = sc.inst[i];
        if (in.hidden) continue;
        float ti = 0.f;
        float di = formula_de_f<true, EXT>(in, instance_local<float>(in, p), &ti) * in.de_factor;
        Mat mi = strip_mat(S.inst[i], ti);
        if (first) { d = di; m = mi; first = false; continue; }
        float nd, h;
        combine_de(in.combine, d, di, in.blend_k, &nd, &h);
        m = mat_mix(m, mi, h);
        d = nd;
    }
    return m;
}

NDD V3<float> vnorm3(const V3<float>& g) {
    float il = __frsqrt_rn(dot3(g, g));
    return mk3(g.x * il, g.y * il, g.z * il);
}
// tetra_gradient / estimate_normal (template.wgsl:1781-1804), e = eps(t)
template <bool EXT>
NDD V3<float> normal_tetra(const DevScene& sc, const V3<float>& p, float e) {
    const float K[4][3] = {{1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {1, 1, 1}};
    V3<float> g = mk3(0.f, 0.f, 0.f);
    for (int i = 0; i < 4; ++i) {
        float d = scene_de<false, EXT>(sc, mk3(p.x + K[i][0] * e, p.y + K[i][1] * e, p.z + K[i][2] * e), nullptr);
        g = mk3(g.x + K[i][0] * d, g.y + K[i][1] * d, g.z + K[i][2] * d);
    }
    return vnorm3(g);
}
// oct_encode / oct_decode (template.wgsl:3696-3718): the octahedral map of a unit normal
// packed as two 16-bit snorm values (pack2x16snorm / unpack2x16snorm)
NDD uint32_t oct_pack(const V3<float>& n) {
    const float l1 = fabsf(n.x) + fabsf(n.y) + fabsf(n.z);
    float px = n.x / l1, py = n.y / l1;
    if (n.z < 0.f) {
        const float qx = (1.f - fabsf(py)) * (px >= 0.f ? 1.f : -1.f);
        const float qy = (1.f - fabsf(px)) * (py >= 0.f ? 1.f : -1.f);
        px = qx; py = qy;
    }
    const int ix = (int)floorf(0.5f + 32767.f * fminf(1.f, fmaxf(-1.f, px)));
    const int iy = (int)floorf(0.5f + 32767.f * fminf(1.f, fmaxf(-1.f, py)));
    return ((uint32_t)ix & 0xffffu) | (((uint32_t)iy & 0xffffu) << 16);
}
NDD V3<float> oct_unpack(uint32_t v) {
    const float fx = fmaxf((float)(int16_t)(v & 0xffffu) / 32767.f, -1.f);
    const float fy = fmaxf((float)(int16_t)(v >> 16) / 32767.f, -1.f);
    float x = fx, y = fy;
    const float z = 1.f - fabsf(fx) - fabsf(fy);
    const float t = fmaxf(-z, 0.f);
    x += x >= 0.f ? -t : t;
    y += y >= 0.f ? -t : t;
    return vnorm3(mk3(x, y, z));
}
// hash33 (template.wgsl:3439-3450): three values in [0, 1) per integer lattice cell
NDD void hash33(float px, float py, float pz, float* out) {
    uint32_t v0 = (uint32_t)(int32_t)px, v1 = (uint32_t)(int32_t)py, v2 = (uint32_t)(int32_t)pz;
    const uint32_t M = 1664525u, A = 1013904223u;
    v0 = v0 * M + A; v1 = v1 * M + A; v2 = v2 * M + A;
    v0 += v1 * v2; v1 += v2 * v0; v2 += v0 * v1;
    v0 ^= v0 >> 16; v1 ^= v1 >> 16; v2 ^= v2 >> 16;
    v0 += v1 * v2; v1 += v2 * v0; v2 += v0 * v1;
    out[0] = (float)v0 * 2.3283064365386963e-10f;
    out[1] = (float)v1 * 2.3283064365386963e-10f;
    out[2] = (float)v2 * 2.3283064365386963e-10f;
}
// film_normal with film_noise (template.wgsl:3452-3476): the normal bent by smooth value
// noise of the position, for the thin-film view angle
NDD V3<float> film_normal(const V3<float>& n, const V3<float>& pos, const Mat& m) {
    const float amount = fclamp(m.film_perturb, 0.f, 1.f);
    if (amount <= 0.001f) return n;  // MATERIAL_EPSILON
    const float is = 1.f / fmaxf(m.film_perturb_scale, 1e-4f);
    const float p[3] = {pos.x * is, pos.y * is, pos.z * is};
    float i[3], w[3];
    for (int k = 0; k < 3; ++k) {
        i[k] = floorf(p[k]);
        const float f = p[k] - i[k];
        w[k] = f * f * (3.f - 2.f * f);
    }
    float acc[3] = {0.f, 0.f, 0.f};
    for (int c = 0; c < 8; ++c) {
        const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
        float h[3];
        hash33(i[0] + (float)dx, i[1] + (float)dy, i[2] + (float)dz, h);
        const float wt = (dx ? w[0] : 1.f - w[0]) * (dy ? w[1] : 1.f - w[1]) * (dz ? w[2] : 1.f - w[2]);
        for (int k = 0; k < 3; ++k) acc[k] += h[k] * wt;
    }
    return vnorm3(mk3(n.x + (acc[0] * 2.f - 1.f) * amount, n.y + (acc[1] * 2.f - 1.f) * amount,
                      n.z + (acc[2] * 2.f - 1.f) * amount));
}
// thin_film_tint (template.wgsl:3478-3519): Airy reflectance of the film over its base per
// wavelength (nm), s and p averaged, divided by its peak over the phase and blended with 1
NDD Sp8 thin_film_tint(float hero, float cos_view, const Mat& m) {
    const float strength = fclamp(m.film_strength, 0.f, 1.f);
    const float theta = acosf(fclamp(cos_view, 0.f, 1.f)) * fmaxf(m.film_angle_scale, 0.f);
    const float c1 = cosf(fminf(theta, 1.5607f));
    const float s1 = sqrtf(fmaxf(1.f - c1 * c1, 0.f));
    const float n2 = fmaxf(m.film_ior, 1.f), n3 = fmaxf(m.ior, 1.f);
    const float s2 = s1 / n2, c2 = sqrtf(fmaxf(1.f - s2 * s2, 0.f));
    const float s3 = s1 / n3, c3 = sqrtf(fmaxf(1.f - s3 * s3, 0.f));
    const float r12s = (c1 - n2 * c2) / (c1 + n2 * c2), r12p = (n2 * c1 - c2) / (n2 * c1 + c2);
    const float r23s = (n2 * c2 - n3 * c3) / (n2 * c2 + n3 * c3), r23p = (n3 * c2 - n2 * c3) / (n3 * c2 + n2 * c3);
    auto peak_of = [](float a, float b) { const float x = (fabsf(a) + fabsf(b)) / (1.f + fabsf(a * b)); return x * x; };
    const float peak = 0.5f * (peak_of(r12s, r23s) + peak_of(r12p, r23p));
    if (peak < 1e-5f) return sp_splat(1.f);  // FILM_MIN_PEAK
    auto airy = [](float a, float b, float cd) { const float x = 2.f * a * b * cd; return (a * a + b * b + x) / (1.f + a * a * b * b + x); };
    const float path = 4.f * 3.14159265358979f * n2 * fmaxf(m.film_thickness, 0.f) * c2;
    float lam[8];
    lambdas_of(hero, lam);
    Sp8 r;
    for (int k = 0; k < 8; ++k) {
        const float cd = cosf(path / lam[k]);
        const float R = 0.5f * (airy(r12s, r23s, cd) + airy(r12p, r23p, cd));
        r.v[k] = 1.f + (R / peak - 1.f) * strength;
    }
    return r;
}
// cosine_hemisphere_from (template.wgsl:3341-3350)
NDD V3<float> cosine_sample(const V3<float>& n, float sx, float sy) {
    const float u1 = hash12w(sx, sy);
    const float u2 = hash12w(add_rn(sx, 91.73f), add_rn(sy, 31.13f));
    const float r = sqrtf(u1);
    float sp, cp;
    sincosf(u2 * 6.28318530718f, &sp, &cp);
    const V3<float> local = mk3(r * cp, r * sp, sqrtf(fmaxf(1.f - u1, 0.f)));
    const V3<float> up = fabsf(n.z) > 0.99f ? mk3(1.f, 0.f, 0.f) : mk3(0.f, 0.f, 1.f);
    const V3<float> tg = vnorm3(mk3(up.y * n.z - up.z * n.y, up.z * n.x - up.x * n.z, up.x * n.y - up.y * n.x));
    const V3<float> bt = mk3(n.y * tg.z - n.z * tg.y, n.z * tg.x - n.x * tg.z, n.x * tg.y - n.y * tg.x);
    return vnorm3(mk3(tg.x * local.x + bt.x * local.y + n.x * local.z, tg.y * local.x + bt.y * local.y + n.y * local.z,
                      tg.z * local.x + bt.z * local.y + n.z * local.z));
}

// light_sample (template.wgsl:2178-2229) at a surface point: point, global, and ray lights
// without a segment to integrate over (a Gaussian beam of radius w = waist + t tan(spread) at
// depth t along its axis, lighting only what lies in front of the source)
NDD void light_of(const McLight& L, float max_dist, const V3<float>& p, V3<float>* dir, V3<float>* col, float* dist) {
    if (L.type == 1) {
        *dir = mk3(L.v[0], L.v[1], L.v[2]);
        *col = mk3(L.color[0], L.color[1], L.color[2]);
        *dist = max_dist * 2.f;
        return;
    }
    V3<float> to = mk3(L.v[0] - p.x, L.v[1] - p.y, L.v[2] - p.z);
    float d = len3(to);
    float inv = 1.f / fmaxf(d, 1e-4f);
    *dir = mk3(to.x * inv, to.y * inv, to.z * inv);
    *dist = d;
    float a;
    if (L.type == 0) {
        a = 1.f / fmaxf(d * d, 0.01f);
    } else {
        const float t0 = -(to.x * L.axis[0] + to.y * L.axis[1] + to.z * L.axis[2]);
        const float qx = -to.x - L.axis[0] * t0, qy = -to.y - L.axis[1] * t0, qz = -to.z - L.axis[2] * t0;
        const float w = L.waist + fmaxf(t0, 0.f) * L.tan_spread;
        const float profile = expf(-(qx * qx + qy * qy + qz * qz) / (w * w));
        a = t0 > 0.f ? profile / (w * w) : 0.f;
    }
    *col = mk3(L.color[0] * a, L.color[1] * a, L.color[2] * a);
}

// ----------------------------------------------------------------------------- path state
struct McPaths {
    // SoA path state
    float *px, *py, *pz, *dx, *dy, *dz;
    float* thr;   // 8 per path
    float* rad;   // 8 per path
    float *hero, *seedx, *seedy;
    float* hit_t;
    uint8_t *depth, *flags;  // flags: 1 reflected, 2 hit, 4 inside a solid
    float* lambda;           // picked wavelength in micrometres (0: none yet)
    uint16_t* inner;         // inside-march step budget of the solid the path is in
    // flags also: 8 the primary segment comes from the cache (no extend), 16 the path's
    // depth-0 event was a mirror reflection (its depth-1 normal is packed like the primary)
    // primary hit cache per pixel (adapt_stats[].primary, template.wgsl:3720-3766): bit 0
    // valid, bit 1 hit; distance; octahedral-packed normal
    float* ct;
    uint32_t *cf, *cn;
    float* dof_ap;  // per pixel: the effective aperture of its centre ray (depth of field)
    // queues (path indices) and counters
    uint32_t *live_a, *live_b;
    uint32_t* shadow_path;
    float *sox, *soy, *soz, *sdx, *sdy, *sdz, *sdist, *sk;
    float* sw;  // 8 per shadow ray
    uint32_t* ind_path;
    float *iox, *ioy, *ioz, *idx, *idy, *idz, *it;
    uint8_t* ihit;
    float* iw;  // 8 per indirect ray
    uint32_t* ctr;  // [0] live in, [1] live out, [2] shadows, [3] indirect, [4..7] queue heads, [8..9] de counters (u64)
};
enum { C_LIVE_IN = 0, C_LIVE_OUT = 1, C_SHADOW = 2, C_IND = 3, C_HEAD_EXT = 4, C_HEAD_SH = 5, C_HEAD_IND = 6,
       C_SEG = 10, C_SHTOT = 11, C_DE = 12 };

// Per-chunk values, read from device memory so that one captured graph of a chunk's
// passes serves every chunk and every call of the same shape.
constexpr uint32_t MC_MAX_CHUNKS = 64;  // chunk records uploaded per batch of chunks
struct McChunk {
    uint32_t sample_base;  // first sample index of the chunk
    uint32_t spp_chunk;    // samples in the chunk
    uint32_t spp_total;    // samples of the call (resolve scales by 1 / spp_total)
    float mc_sample0;      // mc_sample of sample 0 (Uniforms.mc_sample)
};

struct McArgs {
    const DevJob* job;
    const DevScene* scene;
    const McScene* S;
    const McChunk* ch;
    McPaths P;
    uint32_t npix;
    uint32_t chunk;  // claim size of the persistent passes
    float* out_rgb;
    float* out_depth;
    float* out_normal;
    float* adapt;  // per pixel: luminance sum, square sum, count, converged; colour sum (3), samples
};

NDD unsigned long long* de_ctr(const McArgs& a) { return (unsigned long long*)(a.P.ctr + C_DE); }

// Path index p = (sample in chunk * npix + pixel) * slots + lens slot.
NDD uint32_t path_pixel(const McArgs& a, uint32_t p) { return (p / (uint32_t)a.S->slots) % a.npix; }
NDD uint32_t path_sample(const McArgs& a, uint32_t p) { return a.ch->sample_base + (p / (uint32_t)a.S->slots) / a.npix; }
NDD uint32_t path_slot(const McArgs& a, uint32_t p) { return p % (uint32_t)a.S->slots; }

// The lens samples a pixel takes per sample (n = mc_inner_count(DOF_SAMPLES_HQ) when its
// effective aperture exceeds DOF_MIN_APERTURE, else 1).
NDD uint32_t pixel_lens_samples(const McArgs& a, uint32_t pix) {
    return (a.S->dof_on && a.P.dof_ap[pix] > 0.0001f) ? (uint32_t)a.S->dof_n : 1u;
}

NDD uint32_t append(uint32_t* ctr, bool want, uint32_t* slot) {
    const uint64_t m = __builtin_amdgcn_ballot_w64(want);
    if (!m) return 0;
    const int leader = __builtin_ctzll(m);
    uint32_t base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(ctr, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    const unsigned lane = __lane_id();
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    *slot = base + (uint32_t)__popcll(m & lt);
    return (uint32_t)__popcll(m);
}

// The scene records as the kernels read them: through kernel arguments declared __restrict__,
// so the compiler knows no store of the kernel reaches them and keeps their loads scalar
// (s_load into SGPRs) instead of per-lane global loads inside the DE loops.
NDD McArgs restricted(const McArgs& a0, const DevScene* __restrict__ scene, const McScene* __restrict__ mcs) {
    McArgs a = a0;
    a.scene = scene;
    a.S = mcs;
    return a;
}

// --- generate: camera paths of samples [sample_base, sample_base + spp_chunk). With adaptive
// sampling, a sample of a pixel that had converged before this chunk starts no path (the
// sample skip of cs_wf_generate, template.wgsl:5134-5146) and the live queue holds only the
// paths started; the resolve drops the same samples, and those of pixels that converge
// within the chunk.
__global__ void k_mc_generate(McArgs a) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;  // one pixel sample
    const uint32_t n = a.npix * a.ch->spp_chunk;
    if (q >= n) return;
    const DevJob& J = *a.job;
    const McScene& S = *a.S;
    const uint32_t pix = q % a.npix, s = a.ch->sample_base + q / a.npix;
    const int x = (int)(pix % (uint32_t)J.rect_w), y = (int)(pix / (uint32_t)J.rect_w);
    const int gx = J.rect_x + x, gy = J.rect_y + y;
    const float ms = add_rn(a.ch->mc_sample0, (float)s);
    const uint32_t slots = (uint32_t)S.slots;
    const uint32_t lens = pixel_lens_samples(a, pix);
    bool skip = false;
    if (S.adapt_on) {
        const bool on = (float)gx < S.adapt_row_width;
        skip = on && ms > 0.5f && a.adapt[8 * (size_t)pix + 3] > 0.5f;
    }
    const uint32_t n_paths = skip ? 0u : lens;
    if (S.adapt_on || S.dof_on) {
        // the live queue holds only the paths started
        uint32_t slot = 0;
        const uint64_t m = __builtin_amdgcn_ballot_w64(true);
        const int leader = __builtin_ctzll(m);
        // wave-wide exclusive prefix of n_paths
        uint32_t incl = n_paths;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)incl, off);
            if ((int)__lane_id() >= off) incl += o;
        }
        const uint32_t total = (uint32_t)__shfl((int)incl, 63 - __builtin_clzll(m));
        uint32_t base = 0;
        if ((int)__lane_id() == leader && total) base = atomicAdd(a.P.ctr + C_LIVE_IN, total);
        base = (uint32_t)__shfl((int)base, leader);
        slot = base + incl - n_paths;
        for (uint32_t j = 0; j < n_paths; ++j) a.P.live_a[slot + j] = q * slots + j;
    } else {
        a.P.live_a[q] = q;
    }
    if (n_paths == 0) return;
    const V3<float> d = primary_dir(J, gx, gy);
    // base seed = frag + mc_sample * (91.73, 57.31) (fs_main, template.wgsl:4818)
    const float bx = mad_rn(ms, 91.73f, (float)gx + 0.5f);
    const float by = mad_rn(ms, 57.31f, (float)gy + 0.5f);
    V3<float> focus = mk3(0.f, 0.f, 0.f);
    float ap = 0.f;
    if (lens > 1u) {
        // focus_t = focus_distance / max(dot(dir, forward), EPSILON); focus point on the ray
        const float ft = S.focus_distance / fmaxf(d.x * J.fwd[0] + d.y * J.fwd[1] + d.z * J.fwd[2], 1e-4f);
        focus = mk3(J.cam[0] + d.x * ft, J.cam[1] + d.y * ft, J.cam[2] + d.z * ft);
        ap = a.P.dof_ap[pix];
    }
    uint8_t cache_fl = 0;
    float cache_t = 0.f;
    if (S.cache_on && lens == 1u) {
        // a primary segment stored by sample 0 of this call (an earlier chunk) is reused
        const uint32_t c = a.P.cf[pix];
        if (c & 1u) {
            cache_fl = (uint8_t)(8u | ((c & 2u) ? 2u : 0u));
            cache_t = a.P.ct[pix];
        }
    }
    for (uint32_t j = 0; j < n_paths; ++j) {
        const uint32_t p = q * slots + j;
        float sx0 = bx, sy0 = by;
        V3<float> o = mk3(J.cam[0], J.cam[1], J.cam[2]), dir = d;
        uint8_t fl = cache_fl;
        if (lens > 1u) {
            // lens sample: seed + s (91.37, 51.91), sample_unit_disk * aperture on the camera's
            // right / up, aimed at the focus point (template.wgsl:5171-5179, 4709-4713)
            sx0 = add_rn(bx, mul_rn((float)j, 91.37f));
            sy0 = add_rn(by, mul_rn((float)j, 51.91f));
            const float r = sqrtf(hash12(sx0, sy0));
            const float th = hash12w(add_rn(sx0, 17.17f), add_rn(sy0, 71.71f)) * 6.28318530718f;
            const float dxl = r * cosf(th) * ap, dyl = r * sinf(th) * ap;
            o = mk3(J.cam[0] + J.right[0] * dxl + J.up[0] * dyl, J.cam[1] + J.right[1] * dxl + J.up[1] * dyl,
                    J.cam[2] + J.right[2] * dxl + J.up[2] * dyl);
            dir = vnorm3(mk3(focus.x - o.x, focus.y - o.y, focus.z - o.z));
            fl = 32u;  // a lens path: no primary hit cache
        }
        a.P.px[p] = o.x; a.P.py[p] = o.y; a.P.pz[p] = o.z;
        a.P.dx[p] = dir.x; a.P.dy[p] = dir.y; a.P.dz[p] = dir.z;
        for (int k = 0; k < 8; ++k) { a.P.thr[8 * p + k] = 1.f; a.P.rad[8 * p + k] = 0.f; }
        // path_begin (template.wgsl:3819-3839)
        a.P.seedx[p] = mad_rn(ms, 27.31f, sx0);
        a.P.seedy[p] = mad_rn(ms, 45.17f, sy0);
        a.P.hero[p] = mad_rn(ms, 0.61803399f, hash12w(add_rn(sx0, 19.31f), add_rn(sy0, 47.11f)));
        a.P.depth[p] = 0;
        if (fl & 8u) a.P.hit_t[p] = cache_t;
        a.P.flags[p] = fl;
        a.P.lambda[p] = 0.f;
        a.P.inner[p] = 64;  // PATH_DEFAULT_INNER_STEPS
    }
}

// --- depth-of-field probe (cs_wf_generate, template.wgsl:5151-5163): once per call, each
// pixel's centre ray is marched and the effective aperture of its hit distance stored:
// aperture * smoothstep(range / 2, range / 2 + falloff, |t - focus_t|) with falloff
// max(4 aperture, 0.05) (dof_aperture, :4726-4730); a miss counts as max_dist.
struct DofProbeRays {
    const McArgs& a;
    NDD void setup(uint32_t i, V3<float>* o, V3<float>* d) const {
        const DevJob& J = *a.job;
        *o = mk3(J.cam[0], J.cam[1], J.cam[2]);
        *d = primary_dir(J, J.rect_x + (int)(i % (uint32_t)J.rect_w), J.rect_y + (int)(i / (uint32_t)J.rect_w));
    }
    NDD void finish(uint32_t i, bool hit, float t) const {
        const DevJob& J = *a.job;
        const McScene& S = *a.S;
        const V3<float> d = primary_dir(J, J.rect_x + (int)(i % (uint32_t)J.rect_w), J.rect_y + (int)(i / (uint32_t)J.rect_w));
        const float ft = S.focus_distance / fmaxf(d.x * J.fwd[0] + d.y * J.fwd[1] + d.z * J.fwd[2], 1e-4f);
        const float probe_t = hit ? t : S.max_dist;
        const float half = fmaxf(S.focus_range, 0.f) * 0.5f;
        const float falloff = fmaxf(S.aperture * 4.f, 0.05f);
        const float x = fclamp((fabsf(probe_t - ft) - half) / falloff, 0.f, 1.f);
        a.P.dof_ap[i] = S.aperture * (x * x * (3.f - 2.f * x));
    }
    NDD int inside(uint32_t) const { return 0; }
    NDD bool skip(uint32_t) const { return false; }
};

// --- persistent march over a queue of rays. RAY provides setup(i, org, dir), finish(i,
// hit, t) and inside(i): the inside-march step budget of a ray that starts in a solid, 0
// for one outside. Outside rays run march_to with the job's march precision. With INSIDE,
// rays in a solid run march_inside (template.wgsl:3015-3046): steps of max(|DE|, floor)
// with the floor growing by 1 + 6 / steps, until DE > 0.0005 t + 0.00005 from the third
// sample on (exited: a hit at t), t > max_dist or the budget ends (no exit).
template <bool EXT, bool INSIDE, class RAY>
NDD void persistent_march(const McArgs& a, RAY& R, uint32_t n, uint32_t* head) {
    const DevScene& sc = *a.scene;
    const DevJob& J = *a.job;
    const MarchParams mp = march_params(J);
    const bool acc = J.accel != nullptr;
    const AccelView av = accel_view(J);
    AccelCache ac;  // every ray of the call reads the same grid: valid across rays
    accel_cache_reset(ac);
    uint32_t cnext = 0, cend = 0;
    bool work_left = true, active = false, first_claim = true;
    uint32_t item = 0;
    V3<float> org = mk3(0.f, 0.f, 0.f), dir = mk3(0.f, 0.f, 1.f);
    March m;
    march_init(m);
    bool ins = false;  // INSIDE: the lane's ray is in a solid
    int ii = 0, isteps = 0;
    float it = 0.f, flo = 0.f, grow = 1.f;
    const float max_dist = a.S->max_dist;
    unsigned long long de = 0;
    for (;;) {
        if (work_left) {
            uint64_t need = __builtin_amdgcn_ballot_w64(!active);
            if (__popcll(need) >= 8 || need == ~0ull) {
                while (need != 0ull) {
                    if (cnext >= cend) {
                        const uint32_t base = claim_chunk(head, n, a.chunk, first_claim);
                        first_claim = false;
                        if (base >= n) { work_left = false; break; }
                        cnext = base;
                        cend = (uint32_t)__builtin_amdgcn_readfirstlane((int)min(base + a.chunk, n));
                    }
                    const uint32_t avail = cend - cnext;
                    const unsigned lane = __lane_id();
                    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
                    const uint32_t rank = (uint32_t)__popcll(need & lt);
                    const bool mine = ((need >> lane) & 1ull) && rank < avail;
                    if (mine && R.skip(cnext + rank)) {
                        // the ray's segment is already known (primary hit cache)
                    } else if (mine) {
                        item = cnext + rank;
                        R.setup(item, &org, &dir);
                        march_init(m);
                        if (INSIDE) {
                            isteps = R.inside(item);
                            ins = isteps > 0;
                            ii = 0; it = 0.f; flo = 0.004f;  // max(INSIDE_STEP_BASE, INSIDE_MIN_STEP)
                            grow = 1.f + 6.f / (float)max(isteps, 1);
                        }
                        active = true;
                    }
                    cnext += min((uint32_t)__popcll(need), avail);
                    need = __builtin_amdgcn_ballot_w64(((need >> lane) & 1ull) && !mine);
                }
            }
        }
        const uint64_t act = __builtin_amdgcn_ballot_w64(active);
        if (act == 0ull) {
            if (!work_left) break;
            continue;
        }
        float dist = 0.f;
        int st = -1;
        bool used = false;  // the lane evaluated a DE this round (no accelerator skip applied)
        if (active && INSIDE && ins) {
            const V3<float> ps = mk3(org.x + dir.x * it, org.y + dir.y * it, org.z + dir.z * it);
            const float d = scene_de<false, EXT>(sc, ps, nullptr);
            used = true;
            if (ii > 1 && d > 0.0005f * it + 0.00005f) {  // inside_exited
                R.finish(item, true, it);
                active = false;
            } else {
                it += fmaxf(fabsf(d), flo);
                flo *= grow;
                ++ii;
                if (it > max_dist || ii >= isteps) {
                    R.finish(item, false, it);
                    active = false;
                }
            }
        } else if (!EXT && !INSIDE && !acc) {
            // Idle lanes step too, on the ray they last held (or the initial one), as in the
            // G-buffer march: the ray state is updated in place instead of copied in and out
            // of a branch every round. An idle lane's result is discarded and its state set
            // again when it takes a ray; the tuned DEs read no memory by position.
            const float s = march_sample_t(m);
            const V3<float> ps = mk3(org.x + dir.x * s, org.y + dir.y * s, org.z + dir.z * s);
            const int st_any = march_update(m, mp, s, scene_de<false, EXT>(sc, ps, nullptr), &dist);
            if (active) {
                used = true;
                if (st_any != MS_CONT) {
                    R.finish(item, st_any == MS_HIT, dist);
                    active = false;
                }
            }
        } else if (active) {
            const float s = march_sample_t(m);
            const V3<float> ps = mk3(org.x + dir.x * s, org.y + dir.y * s, org.z + dir.z * s);
            // an accelerator skip takes the step without a DE (template.wgsl:1890-1900)
            if (acc && m.probing) st = march_skip_step(m, mp, accel_skip_cached(av, ac, ps), &dist);
            used = st < 0;
            if (used) st = march_update(m, mp, s, scene_de<false, EXT>(sc, ps, nullptr), &dist);
            if (st != MS_CONT) {
                R.finish(item, st == MS_HIT, dist);
                active = false;
            }
        }
        de += (unsigned long long)__popcll(__builtin_amdgcn_ballot_w64(used));
    }
    if (__lane_id() == 0 && de) atomicAdd(de_ctr(a), de);
}

struct ExtendRays {
    const McArgs& a;
    const uint32_t* q;
    NDD void setup(uint32_t i, V3<float>* o, V3<float>* d) const {
        const uint32_t p = q[i];
        *o = mk3(a.P.px[p], a.P.py[p], a.P.pz[p]);
        *d = mk3(a.P.dx[p], a.P.dy[p], a.P.dz[p]);
    }
    NDD void finish(uint32_t i, bool hit, float t) const {
        const uint32_t p = q[i];
        a.P.hit_t[p] = t;
        a.P.flags[p] = (uint8_t)((a.P.flags[p] & 53u) | (hit ? 2u : 0u));
    }
    NDD bool skip(uint32_t i) const { return (a.P.flags[q[i]] & 8u) != 0; }
    NDD int inside(uint32_t i) const {
        const uint32_t p = q[i];
        return (a.P.flags[p] & 4u) ? (int)a.P.inner[p] : 0;
    }
};
struct IndirectRays {
    const McArgs& a;
    NDD void setup(uint32_t i, V3<float>* o, V3<float>* d) const {
        *o = mk3(a.P.iox[i], a.P.ioy[i], a.P.ioz[i]);
        *d = mk3(a.P.idx[i], a.P.idy[i], a.P.idz[i]);
    }
    NDD void finish(uint32_t i, bool hit, float t) const {
        a.P.it[i] = t;
        a.P.ihit[i] = hit ? 1 : 0;
    }
    NDD int inside(uint32_t) const { return 0; }
    NDD bool skip(uint32_t) const { return false; }
};

// The two march kernels of the path trace run at 7 waves per SIMD (72 VGPRs, five values
// spilled) instead of the compiler's 6: a 64-spp frame takes 2.7% less time. The extend
// kernel's INSIDE variant (scenes with transmissive materials) keeps the compiler's choice.
template <bool EXT, bool INSIDE>
__global__ void __launch_bounds__(64)
__attribute__((amdgpu_waves_per_eu((EXT || INSIDE) ? 1 : 7, (EXT || INSIDE) ? 8 : 7))) k_mc_extend(McArgs a0, int ping, const DevScene* __restrict__ scene, const McScene* __restrict__ mcs) {
    McArgs a = restricted(a0, scene, mcs);
    const uint32_t* q = ping ? a.P.live_b : a.P.live_a;
    const uint32_t n = a.P.ctr[C_LIVE_IN];
    ExtendRays R{a, q};
    persistent_march<EXT, INSIDE>(a, R, n, a.P.ctr + C_HEAD_EXT);
}

template <bool EXT>
__global__ void __launch_bounds__(64) k_mc_dof_probe(McArgs a0, const DevScene* __restrict__ scene, const McScene* __restrict__ mcs) {
    McArgs a = restricted(a0, scene, mcs);
    DofProbeRays R{a};
    persistent_march<EXT, false>(a, R, a.npix, a.P.ctr + C_HEAD_EXT);
}

template <bool EXT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(EXT ? 1 : 7, EXT ? 8 : 7))) k_mc_indirect_march(McArgs a0, const DevScene* __restrict__ scene, const McScene* __restrict__ mcs) {
    McArgs a = restricted(a0, scene, mcs);
    IndirectRays R{a};
    persistent_march<EXT, false>(a, R, a.P.ctr[C_IND], a.P.ctr + C_HEAD_IND);
}

// shade_local (template.wgsl:3222-3331) for subsurface-free materials: the albedo and the
// specular term carry the thin-film tint (film).
// With defer, each shadowed light's contribution is queued as a shadow ray with weight
// w_scale * light * (albedo n.l + specular); otherwise it is added unshadowed. With
// pick_one (several lights, mc_light_pick), the shadowed lights are not queued each: one is
// chosen with probability weight / total, weight = (r + g + b) (n.l + specular), from
// hash12w(seed + (47.23, 113.91)), and queued scaled by total / weight.
NDD void queue_shadow(const McArgs& a, uint32_t path, const V3<float>& pos, const V3<float>& n, const V3<float>& ld,
                      float dist, float k, const Sp8& c, const Sp8& w_scale, float scale) {
    uint32_t slot;
    append(a.P.ctr + C_SHADOW, true, &slot);
    a.P.shadow_path[slot] = path;
    a.P.sox[slot] = pos.x + n.x * 0.004f;
    a.P.soy[slot] = pos.y + n.y * 0.004f;
    a.P.soz[slot] = pos.z + n.z * 0.004f;
    a.P.sdx[slot] = ld.x; a.P.sdy[slot] = ld.y; a.P.sdz[slot] = ld.z;
    a.P.sdist[slot] = dist;
    a.P.sk[slot] = k;
    for (int q = 0; q < 8; ++q) a.P.sw[8 * slot + q] = c.v[q] * w_scale.v[q] * scale;
}
// film_albedo: the material's albedo times its thin-film tint at this view (g_shade_albedo).
NDD Sp8 film_of(float hero, const V3<float>& pos, const V3<float>& dir, const V3<float>& n, const Mat& m) {
    if (!(m.film_strength > 0.001f)) return sp_splat(1.f);  // MATERIAL_EPSILON
    const V3<float> fn = film_normal(n, pos, m);
    return thin_film_tint(hero, -(dir.x * fn.x + dir.y * fn.y + dir.z * fn.z), m);
}
NDD Sp8 shade_lights(const McArgs& a, const Basis& B, const V3<float>& pos, const V3<float>& dir, const V3<float>& n,
                     const Mat& m, const Sp8& film, const Sp8& ambient, bool shadows, const Sp8& w_scale, uint32_t path,
                     float sx, float sy) {
    const McScene& S = *a.S;
    const V3<float> view = mk3(-dir.x, -dir.y, -dir.z);
    const float spec_power = 4.f + 124.f * fclamp(m.glossiness, 0.f, 1.f);
    Sp8 albedo = from_rgb(B, m.color[0], m.color[1], m.color[2]);
    for (int k = 0; k < 8; ++k) albedo.v[k] *= film.v[k];
    const bool pick_one = shadows && S.light_pick && S.light_count > 1;
    // light i's direction, colour, distance, n.l and specular term (light_terms)
    auto terms = [&](int i, V3<float>* ld, V3<float>* col, float* dist, float* ndotl, float* spec) {
        const McLight& L = S.lights[i];
        light_of(L, S.max_dist, pos, ld, col, dist);
        *ndotl = fmaxf(dot3(n, *ld), 0.f);
        const V3<float> hd = vnorm3(mk3(ld->x + view.x, ld->y + view.y, ld->z + view.z));
        *spec = S.specular_on ? powf(fmaxf(dot3(n, hd), 0.f), spec_power) * m.glossiness : 0.f;
    };
    Sp8 acc = sp_splat(0.f);
    float pick_total = 0.f;
    for (int i = 0; i < S.light_count; ++i) {
        const McLight& L = S.lights[i];
        V3<float> ld, col;
        float dist, ndotl, spec;
        terms(i, &ld, &col, &dist, &ndotl, &spec);
        if (fmaxf(col.x, fmaxf(col.y, col.z)) < 1e-5f) continue;
        if (ndotl <= 0.f) continue;
        if (pick_one && L.cast) {
            pick_total += (col.x + col.y + col.z) * (ndotl + spec);
            continue;
        }
        const Sp8 ls = from_rgb(B, col.x, col.y, col.z);
        Sp8 c;
        for (int k = 0; k < 8; ++k) c.v[k] = ls.v[k] * (albedo.v[k] * ndotl + spec * film.v[k]);
        if (shadows && L.cast && S.shadow_steps > 0) {
            queue_shadow(a, path, pos, n, ld, dist, L.k, c, w_scale, 1.f);
        } else {
            for (int k = 0; k < 8; ++k) acc.v[k] += c.v[k];
        }
    }
    if (pick_total > 0.f) {
        const float at = hash12w(add_rn(sx, 47.23f), add_rn(sy, 113.91f)) * pick_total;
        int pick = 0;
        float running = 0.f, pick_w = 0.f;
        for (int i = 0; i < S.light_count; ++i) {
            const McLight& L = S.lights[i];
            if (!L.cast) continue;
            V3<float> ld, col;
            float dist, ndotl, spec;
            terms(i, &ld, &col, &dist, &ndotl, &spec);
            if (fmaxf(col.x, fmaxf(col.y, col.z)) < 1e-5f || ndotl <= 0.f) continue;
            const float w = (col.x + col.y + col.z) * (ndotl + spec);
            if (w <= 0.f) continue;
            pick = i;
            pick_w = w;
            running += w;
            if (running > at) break;
        }
        V3<float> ld, col;
        float dist, ndotl, spec;
        terms(pick, &ld, &col, &dist, &ndotl, &spec);
        const Sp8 ls = from_rgb(B, col.x, col.y, col.z);
        Sp8 c;
        for (int k = 0; k < 8; ++k) c.v[k] = ls.v[k] * (albedo.v[k] * ndotl + spec * film.v[k]);
        const float scale = pick_total / pick_w;
        if (S.shadow_steps > 0) {
            queue_shadow(a, path, pos, n, ld, dist, S.lights[pick].k, c, w_scale, scale);
        } else {
            for (int k = 0; k < 8; ++k) acc.v[k] += c.v[k] * scale;
        }
    }
    Sp8 r;
    for (int k = 0; k < 8; ++k) r.v[k] = albedo.v[k] * ambient.v[k] + acc.v[k];
    return r;
}

// ----------------------------------------------------------------------------- subsurface
// Subsurface / subsurface_setup (template.wgsl:3088-3114): amount, radius (at least 0.0005),
// each wavelength's reach = max((albedo / peak)^2, 0.05), and the blurred normal: the
// tetrahedral gradient at pos + n radius/2 with taps radius/4.
struct Sss {
    float amount, radius;
    Sp8 reach;
    V3<float> blur;
};
template <bool EXT>
NDD Sss sss_setup(const DevScene& sc, const V3<float>& pos, const V3<float>& n, const Mat& m, const Sp8& albedo,
                  unsigned long long& de) {
    Sss s;
    s.amount = fclamp(m.subsurface, 0.f, 1.f);
    s.radius = fmaxf(m.subsurface_radius, 0.0005f);  // SSS_MIN_RADIUS
    s.blur = n;
    s.reach = sp_splat(0.f);
    if (s.amount <= 0.001f) {  // MATERIAL_EPSILON
        s.amount = 0.f;
        return s;
    }
    const float peak = fmaxf(sp_max(albedo), 1e-4f);  // EPSILON
    for (int k = 0; k < 8; ++k) {
        const float rel = albedo.v[k] / peak;
        s.reach.v[k] = fmaxf(rel * rel, 0.05f);  // SSS_REACH_FLOOR
    }
    const float blur = s.radius * 0.5f;  // SSS_BLUR_PER_RADIUS
    const float e = blur * 0.5f;         // SSS_BLUR_TAP_FRACTION
    const V3<float> c = mk3(pos.x + n.x * blur, pos.y + n.y * blur, pos.z + n.z * blur);
    const float K[4][3] = {{1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {1, 1, 1}};
    V3<float> g = mk3(0.f, 0.f, 0.f);
    for (int i = 0; i < 4; ++i) {
        const float d = scene_de<false, EXT>(sc, mk3(c.x + K[i][0] * e, c.y + K[i][1] * e, c.z + K[i][2] * e), nullptr);
        g = mk3(g.x + K[i][0] * d, g.y + K[i][1] * d, g.z + K[i][2] * d);
    }
    de += 4;
    const float len = sqrtf(dot3(g, g));
    if (len > 1e-12f) s.blur = mk3(g.x / len, g.y / len, g.z / len);  // EPSILON_TINY
    return s;
}
// diffuse_weight + subsurface_wrap (template.wgsl:3116-3134): n.l blended toward the blurred
// normal's by the reach, wrapped by 0.75 reach (SSS_WRAP), plus light from behind
NDD Sp8 sss_weight(const Sss& s, float ndl, const V3<float>& ld) {
    const float nb = dot3(s.blur, ld);
    Sp8 r;
    for (int k = 0; k < 8; ++k) {
        const float re = s.reach.v[k];
        const float w = 0.75f * re;
        const float wrap = fmaxf(ndl * (1.f - re) + nb * re + w, 0.f) / (1.f + w);
        r.v[k] = (1.f - s.amount) * fmaxf(ndl, 0.f) + s.amount * (wrap + fmaxf(-ndl, 0.f));
    }
    return r;
}
// subsurface_transmit (template.wgsl:3136-3138)
NDD Sp8 sss_transmit(const Sss& s, float depth) {
    Sp8 r;
    for (int k = 0; k < 8; ++k) r.v[k] = expf(-depth / (s.radius * s.reach.v[k]));
    return r;
}
// shadow_factor + visibility_march for one ray, in line (template.wgsl:2282-2342, 2365-2375):
// the same march as k_mc_shadow, for the rest of a subsurface light path past a surface
// without subsurface
template <bool EXT>
NDD Sp8 vis_inline(const McArgs& a, const Basis& B, const V3<float>& o, const V3<float>& ld, float ldist, float k,
                   unsigned long long& de) {
    const DevScene& sc = *a.scene;
    const McScene& S = *a.S;
    const DevJob& J = *a.job;
    const bool acc = J.accel != nullptr;
    const AccelView av = accel_view(J);
    AccelCache ac;
    accel_cache_reset(ac);
    float t = 0.004f, res = 1.f, flo = 0.004f;  // SHADOW_BIAS as t0, SHADOW_STEP_FLOOR
    bool ins = false;
    Sp8 tint = sp_splat(1.f);
    for (int step = 0; step < S.shadow_steps; ++step) {
        if (t >= ldist) break;
        const V3<float> ps = mk3(o.x + ld.x * t, o.y + ld.y * t, o.z + ld.z * t);
        if (!ins && acc) {
            const float skip = accel_skip_cached(av, ac, ps);
            if (shadow_skip_applies(skip, k, res, t)) {
                t += skip;
                continue;
            }
        }
        const float d = scene_de<false, EXT>(sc, ps, nullptr);
        de++;
        if (ins) {
            if (d > 0.001f * t + 0.0005f) ins = false;  // SHADOW_EXIT_SLOPE, _BIAS
            t += fmaxf(fabsf(d), flo);
            flo *= 1.15f;  // SHADOW_FLOOR_GROWTH
            continue;
        }
        if (d < 0.0005f) {  // SHADOW_HIT_EPS
            if (!S.transp) return sp_splat(0.f);
            const Mat om = hit_material<EXT>(sc, S, ps);
            de++;
            const float al = fclamp(om.transparency, 0.f, 1.f);
            if (al < 0.01f) return sp_splat(0.f);  // SHADOW_OPAQUE_ALPHA
            const Sp8 tc = from_rgb(B, 1.f + (om.color[0] - 1.f) * al, 1.f + (om.color[1] - 1.f) * al,
                                    1.f + (om.color[2] - 1.f) * al);
            for (int q = 0; q < 8; ++q) tint.v[q] *= tc.v[q] * al;
            if (sp_max(tint) < 0.004f) return sp_splat(0.f);  // SHADOW_TINT_CUT
            ins = true;
            t += flo;
            continue;
        }
        res = fminf(res, k * d / t);
        t += fmaxf(d, 0.001f);  // SHADOW_MIN_STEP
    }
    const float vis = fclamp(res, 0.f, 1.f);
    for (int q = 0; q < 8; ++q) tint.v[q] *= vis;
    return tint;
}
// subsurface_light_path (template.wgsl:3140-3218): the transmission toward a light through
// subsurface material. Inside, steps of max(|DE|, floor) with the floor from
// max(0.05 radius, 0.0005) growing by 1 + 6/64; the exit (DE > 0.0005 t + 0.00005 after the
// warm-up) is refined by 7 bisections; more than 6 radii inside gives 0. Outside, steps of
// max(DE, 0.001) with accelerator skips; a surface without subsurface hands over to the
// soft shadow march. Without occlusion the march ends at the first exit.
template <bool EXT>
NDD Sp8 sss_light_path(const McArgs& a, const Basis& B, const V3<float>& pos, const V3<float>& n, bool start_inside,
                       const Sss& s, const McLight& L, const V3<float>& ld, float ldist, bool occlusion,
                       unsigned long long& de) {
    if (!start_inside && !occlusion) return sp_splat(1.f);
    const DevScene& sc = *a.scene;
    const McScene& S = *a.S;
    const DevJob& J = *a.job;
    const bool acc = J.accel != nullptr;
    const AccelView av = accel_view(J);
    AccelCache ac;
    accel_cache_reset(ac);
    const V3<float> o = start_inside ? pos : mk3(pos.x + n.x * 0.004f, pos.y + n.y * 0.004f, pos.z + n.z * 0.004f);
    const float cap = s.radius * 6.f;                          // SSS_PROBE_DEPTH_RADII
    const float step_base = fmaxf(s.radius * 0.05f, 0.0005f);  // SSS_PROBE_STEP_PER_RADIUS, INSIDE_MIN_STEP
    const float growth = 1.f + 6.f / 64.f;                     // INSIDE_STEP_GROWTH / SSS_MAX_STEPS
    bool inside = start_inside;
    float t = 0.f, prev = 0.f, entry = 0.f, depth = 0.f, flo = step_base;
    int inner = 0;
    for (int i = 0; i < S.sss_steps; ++i) {
        if (t >= ldist) break;
        const V3<float> p = mk3(o.x + ld.x * t, o.y + ld.y * t, o.z + ld.z * t);
        if (!inside && acc) {
            const float skip = accel_skip_cached(av, ac, p);
            if (skip > 0.001f) {  // SHADOW_MIN_STEP
                t += skip;
                continue;
            }
        }
        const float d = scene_de<false, EXT>(sc, p, nullptr);
        de++;
        if (inside) {
            if (inner > 1 && d > 0.0005f * t + 0.00005f) {  // INSIDE_WARMUP_STEPS, inside_exited
                float lo = prev, hi = t;                     // refine_inside_exit, 7 steps
                for (int r = 0; r < 7; ++r) {
                    const float mid = 0.5f * (lo + hi);
                    const float dm = scene_de<false, EXT>(sc, mk3(o.x + ld.x * mid, o.y + ld.y * mid, o.z + ld.z * mid), nullptr);
                    if (dm > 0.0005f * mid + 0.00005f) hi = mid;
                    else lo = mid;
                }
                de += 7;
                depth += hi - entry;
                inside = false;
                if (!occlusion) break;
                t = hi + 0.004f;  // SHADOW_BIAS
                continue;
            }
            prev = t;
            t += fmaxf(fabsf(d), flo);
            flo *= growth;
            inner++;
            if (depth + t - entry > cap) return sp_splat(0.f);
            continue;
        }
        if (d < 0.0005f) {  // SHADOW_HIT_EPS
            const Mat hm = part_mask(hit_material<EXT>(sc, S, p), S);
            de++;
            if (!(hm.subsurface > 0.001f)) {
                Sp8 rest = sp_splat(1.f);
                if (L.cast) {
                    // shadow_factor(p - 2 bias dir, dir, light, dir, ldist - t + 2 bias)
                    const V3<float> sp = mk3(p.x - ld.x * 0.008f, p.y - ld.y * 0.008f, p.z - ld.z * 0.008f);
                    rest = vis_inline<EXT>(a, B, mk3(sp.x + ld.x * 0.004f, sp.y + ld.y * 0.004f, sp.z + ld.z * 0.004f),
                                           ld, ldist - t + 0.008f, L.k, de);
                }
                const Sp8 tr = sss_transmit(s, depth);
                for (int q = 0; q < 8; ++q) rest.v[q] *= tr.v[q];
                return rest;
            }
            inside = true;
            entry = t;
            prev = t;
            inner = 1;
            flo = step_base;
            t += flo;
            continue;
        }
        t += fmaxf(d, 0.001f);
    }
    if (inside) depth += fminf(t, ldist) - entry;
    return sss_transmit(s, depth);
}
// shade_local for a subsurface material (template.wgsl:3222-3331, the sss.amount > 0 branch):
// every light is evaluated here through subsurface_light_path (nothing is queued and no
// light is picked); diffuse weight from diffuse_weight, specular only where n.l > 0.
template <bool EXT>
NDD Sp8 shade_lights_sss(const McArgs& a, const Basis& B, const V3<float>& pos, const V3<float>& dir,
                         const V3<float>& n, const Mat& m, const Sp8& film, const Sp8& ambient, bool shadows,
                         unsigned long long& de) {
    const McScene& S = *a.S;
    const V3<float> view = mk3(-dir.x, -dir.y, -dir.z);
    const float spec_power = 4.f + 124.f * fclamp(m.glossiness, 0.f, 1.f);
    Sp8 albedo = from_rgb(B, m.color[0], m.color[1], m.color[2]);
    for (int k = 0; k < 8; ++k) albedo.v[k] *= film.v[k];
    const Sss s = sss_setup<EXT>(*a.scene, pos, n, m, albedo, de);
    Sp8 acc = sp_splat(0.f);
    for (int i = 0; i < S.light_count; ++i) {
        const McLight& L = S.lights[i];
        V3<float> ld, col;
        float dist;
        light_of(L, S.max_dist, pos, &ld, &col, &dist);
        if (fmaxf(col.x, fmaxf(col.y, col.z)) < 1e-5f) continue;  // SURFACE_LIGHT_CUTOFF
        const float ndl = dot3(n, ld);
        const V3<float> hd = vnorm3(mk3(ld.x + view.x, ld.y + view.y, ld.z + view.z));
        const float spec = S.specular_on ? powf(fmaxf(dot3(n, hd), 0.f), spec_power) * m.glossiness : 0.f;
        const Sp8 w = sss_weight(s, ndl, ld);
        const float sspec = ndl > 0.f ? spec : 0.f;
        if (sp_max(w) <= 0.f && sspec <= 0.f) continue;
        const Sp8 ls = from_rgb(B, col.x, col.y, col.z);
        const Sp8 tr = sss_light_path<EXT>(a, B, pos, n, ndl <= 0.f, s, L, ld, dist, shadows && L.cast, de);
        for (int k = 0; k < 8; ++k) acc.v[k] += ls.v[k] * tr.v[k] * (albedo.v[k] * w.v[k] + sspec * film.v[k]);
    }
    Sp8 r;
    for (int k = 0; k < 8; ++k) r.v[k] = albedo.v[k] * ambient.v[k] + acc.v[k];
    return r;
}

// refract (WGSL builtin) and refract_safe (template.wgsl:3387-3394): the normal is turned to
// face the ray; total internal reflection returns the mirror direction.
NDD V3<float> refract_safe(const V3<float>& d, const V3<float>& raw_n, float eta, bool* tir) {
    const V3<float> n = dot3(d, raw_n) > 0.f ? mk3(-raw_n.x, -raw_n.y, -raw_n.z) : raw_n;
    const float c = dot3(n, d);
    const float k = 1.f - eta * eta * (1.f - c * c);
    V3<float> r = mk3(0.f, 0.f, 0.f);
    if (k >= 0.f) {
        const float s = eta * c + sqrtf(k);
        r = mk3(d.x * eta - n.x * s, d.y * eta - n.y * s, d.z * eta - n.z * s);
    }
    *tir = !(dot3(r, r) > 1e-4f);  // REFRACT_TIR_EPSILON
    if (*tir) {
        const float dn = 2.f * c;
        r = mk3(d.x - n.x * dn, d.y - n.y * dn, d.z - n.z * dn);
    }
    return r;
}
// cauchy_ior (template.wgsl:3403-3409), lambda in micrometres
NDD float cauchy_ior(float ior_d, float abbe, float lambda) {
    const float b = (ior_d - 1.f) / (fmaxf(abbe, 1.f) * 1.9099f);
    return ior_d + b * (1.f / (lambda * lambda) - 2.89626f);
}

// --- shade (path_shade, template.wgsl:3902-4124). One path per lane. With SSS, a hit on a
// subsurface material evaluates its lights in line through the material.
template <bool EXT, bool SSS>
__global__ void __launch_bounds__(64) k_mc_shade(McArgs a0, int ping, const DevScene* __restrict__ scene, const McScene* __restrict__ mcs) {
    McArgs a = restricted(a0, scene, mcs);
    const uint32_t* qin = ping ? a.P.live_b : a.P.live_a;
    uint32_t* qout = ping ? a.P.live_a : a.P.live_b;
    const uint32_t n = a.P.ctr[C_LIVE_IN];
    const DevJob& J = *a.job;
    const DevScene& sc = *a.scene;
    const McScene& S = *a.S;
    unsigned long long de = 0;
    for (uint32_t i = blockIdx.x * 64 + threadIdx.x; i - threadIdx.x < n; i += gridDim.x * 64) {
        bool cont = false;
        uint32_t p = 0;
        if (i < n) {
            p = qin[i];
            const int depth = a.P.depth[p];
            const uint8_t fl = a.P.flags[p];
            const bool reflected = fl & 1u;
            const bool in_solid = fl & 4u;
            const V3<float> pos = mk3(a.P.px[p], a.P.py[p], a.P.pz[p]);
            const V3<float> dir = mk3(a.P.dx[p], a.P.dy[p], a.P.dz[p]);
            const float hero = a.P.hero[p];
            const Basis B = basis_of(hero);
            Sp8 thr, rad;
            for (int k = 0; k < 8; ++k) { thr.v[k] = a.P.thr[8 * p + k]; rad.v[k] = a.P.rad[8 * p + k]; }
            // path_seed (template.wgsl:3845-3847)
            const float sx = add_rn(a.P.seedx[p], mad_rn((float)depth, 17.77f, 3.1f));
            const float sy = add_rn(a.P.seedy[p], mad_rn((float)depth, 41.13f, 7.7f));
            const uint32_t pix = path_pixel(a, p);
            // the pixel's own outputs and cache entry come from sample 0 (its first lens slot)
            const bool first = path_sample(a, p) == 0 && path_slot(a, p) == 0;
            const bool lens = (fl & 32u) != 0;
            if (!(fl & 2u)) {
                const V3<float> sky = sky_color(S, dir);
                const Sp8 sk = from_rgb(B, sky.x, sky.y, sky.z);
                for (int k = 0; k < 8; ++k) rad.v[k] += thr.v[k] * sk.v[k];
                if (depth == 0 && first) {
                    if (a.out_depth) a.out_depth[pix] = 0.f;
                    if (a.out_normal) { a.out_normal[3 * pix] = 0.f; a.out_normal[3 * pix + 1] = 0.f; a.out_normal[3 * pix + 2] = 0.f; }
                    if (S.cache_on && !lens) { a.P.ct[pix] = a.P.hit_t[p]; a.P.cf[pix] = 1u; }  // a cached miss
                }
            } else if (sp_max(thr) >= 0.002f) {
                const float t = a.P.hit_t[p];
                const float eps = J.eps_c * t + J.eps_f;
                const V3<float> hp = mk3(pos.x + dir.x * t, pos.y + dir.y * t, pos.z + dir.z * t);
                // path_normal (template.wgsl:3884-3896): with the primary hit cache on, the
                // primary normal and the normal after a depth-0 mirror are octahedral-packed;
                // sample 0 stores the primary hit, later samples read it back
                V3<float> nrm;
                if (fl & 8u) {
                    nrm = oct_unpack(a.P.cn[pix]);
                } else {
                    nrm = normal_tetra<EXT>(sc, hp, eps);
                    de += 4;
                    if (S.cache_on && !lens && !in_solid && (depth == 0 || (depth == 1 && (fl & 16u)))) {
                        const uint32_t packed = oct_pack(nrm);
                        nrm = oct_unpack(packed);
                        if (depth == 0 && first) {
                            a.P.cn[pix] = packed;
                            a.P.ct[pix] = t;
                            a.P.cf[pix] = 3u;  // valid, hit
                        }
                    }
                }
                const Mat mat = part_mask(hit_material<EXT>(sc, S, hp), S);
                de += 1;
                if (depth == 0 && first) {
                    if (a.out_depth) a.out_depth[pix] = t;
                    if (a.out_normal) { a.out_normal[3 * pix] = nrm.x; a.out_normal[3 * pix + 1] = nrm.y; a.out_normal[3 * pix + 2] = nrm.z; }
                }
                // dispersion: the first dispersive hit picks one of the 8 wavelengths
                // (template.wgsl:3995-4003); without soft weights only that lane carries on
                float lambda = a.P.lambda[p];
                int collapse = -1;
                if (lambda == 0.f && mat.abbe > 0.5f) {  // MIN_ABBE
                    const float hx = hash12w(add_rn(sx, 3.71f), add_rn(sy, 8.13f));  // DISPERSION_PICK_SEED_OFFSET
                    const int pick = min((int)(hx * 8.f), 7);
                    float lam[8];
                    lambdas_of(hero, lam);
                    lambda = lam[pick] * 0.001f;
                    a.P.lambda[p] = lambda;
                    if (S.disp_soft <= 0.f) collapse = pick;
                }
                // scatter_at with allow_diffuse = false (template.wgsl:3552-3616)
                const V3<float> nf = dot3(dir, nrm) > 0.f ? mk3(-nrm.x, -nrm.y, -nrm.z) : nrm;
                float ior = fmaxf(mat.ior, 1.f);
                if (lambda > 0.f && mat.abbe > 0.5f) ior = fmaxf(cauchy_ior(ior, mat.abbe, lambda), 1.f);
                const float eta = in_solid ? ior : 1.f / ior;
                const float rough = fclamp(mat.roughness, 0.f, 1.f);
                const float alpha = rough * rough;
                V3<float> nr = nf;
                if (alpha > 1e-4f) {
                    const V3<float> hs = cosine_sample(nf, add_rn(sx, 5.19f), add_rn(sy, 12.73f));
                    nr = vnorm3(mk3(nf.x + (hs.x - nf.x) * alpha, nf.y + (hs.y - nf.y) * alpha, nf.z + (hs.z - nf.z) * alpha));
                }
                const float r0r = (1.f - ior) / (1.f + ior);
                const float r0 = r0r * r0r;
                const float cth = fclamp(1.f - fabsf(dot3(dir, nr)), 0.f, 1.f);
                const float c2 = cth * cth;
                const float f = r0 + (1.f - r0) * c2 * c2 * cth;
                const float transparency = fclamp(mat.transparency, 0.f, 1.f);
                const float refl = fclamp(mat.reflectiveness, 0.f, 1.f);
                const float p_reflect = fclamp(refl + (1.f - refl) * f, 0.f, 1.f);
                const float rest = 1.f - p_reflect;
                const float p_transmit = rest * transparency;
                const float local_share = rest * (1.f - transparency);
                const float p_cont = p_reflect + p_transmit;
                if (local_share > 0.001f) {
                    Sp8 ambient = sp_splat(0.15f);
                    Sp8 wscale;
                    for (int k = 0; k < 8; ++k) wscale.v[k] = thr.v[k] * local_share;
                    const Sp8 film = film_of(hero, hp, dir, nrm, mat);
                    if (depth == 0 && S.indirect_on) {
                        // deferred indirect_light: weight throughput * local_share * albedo
                        // (with its film tint, g_shade_albedo)
                        Sp8 albedo = from_rgb(B, mat.color[0], mat.color[1], mat.color[2]);
                        for (int k = 0; k < 8; ++k) albedo.v[k] *= film.v[k];
                        const V3<float> idir = cosine_sample(nrm, add_rn(sx, 13.13f), add_rn(sy, 71.71f));
                        uint32_t slot;
                        append(a.P.ctr + C_IND, true, &slot);
                        a.P.ind_path[slot] = p;
                        a.P.iox[slot] = hp.x + nrm.x * 0.004f;
                        a.P.ioy[slot] = hp.y + nrm.y * 0.004f;
                        a.P.ioz[slot] = hp.z + nrm.z * 0.004f;
                        a.P.idx[slot] = idir.x; a.P.idy[slot] = idir.y; a.P.idz[slot] = idir.z;
                        for (int k = 0; k < 8; ++k) a.P.iw[8 * slot + k] = wscale.v[k] * albedo.v[k];
                        ambient = sp_splat(0.f);
                    } else if (!reflected && S.ao_on) {
                        // calc_ao (template.wgsl:1825-1841)
                        const float radius = fclamp(t * 0.04f, 0.003f, 0.25f);
                        float occ = 0.f, w = 1.f;
                        for (int k = 0; k < 5; ++k) {
                            const float step = radius * (0.2f + 0.2f * (float)k);
                            const float dd = scene_de<false, EXT>(sc, mk3(hp.x + nrm.x * step, hp.y + nrm.y * step, hp.z + nrm.z * step), nullptr);
                            occ += fmaxf(step - dd, 0.f) * w;
                            w *= 0.6f;
                        }
                        de += 5;
                        ambient = sp_splat(0.15f * fclamp(1.f - 2.f * occ / radius, 0.f, 1.f));
                    }
                    const Sp8 lit = (SSS && mat.subsurface > 0.001f)
                                        ? shade_lights_sss<EXT>(a, B, hp, dir, nrm, mat, film, ambient, !reflected && S.shadows_on, de)
                                        : shade_lights(a, B, hp, dir, nrm, mat, film, ambient, !reflected && S.shadows_on, wscale, p, sx, sy);
                    for (int k = 0; k < 8; ++k) rad.v[k] += wscale.v[k] * lit.v[k];
                }
                // the continuation: r = hash * p_cont picks reflection below p_reflect, else
                // refraction; nothing is left for r >= p_cont (absorbed)
                const float r = hash12w(sx, sy) * p_cont;
                int kind = 0;  // 0 absorbed, 1 reflect, 2 transmit
                V3<float> nd = dir;
                float wrgb[3] = {0.f, 0.f, 0.f};
                bool soft_w = false;
                bool film_w = false;
                if (r < p_reflect) {
                    const float dn = 2.f * dot3(nr, dir);
                    nd = mk3(dir.x - nr.x * dn, dir.y - nr.y * dn, dir.z - nr.z * dn);
                    for (int k = 0; k < 3; ++k) wrgb[k] = mat.color[k] + (1.f - mat.color[k]) * f;
                    film_w = !in_solid;  // the reflection carries the film tint (scatter_at)
                    kind = 1;
                } else if (r < p_reflect + p_transmit) {
                    bool tir;
                    nd = refract_safe(dir, nr, eta, &tir);
                    soft_w = S.disp_soft > 0.f && lambda > 0.f && mat.abbe > 0.5f;
                    for (int k = 0; k < 3; ++k) wrgb[k] = mat.color[k];
                    kind = tir ? 1 : 2;
                }
                if (kind != 0 && !(kind == 1 && !in_solid && !S.reflections_on)) {
                    Sp8 w = from_rgb(B, wrgb[0], wrgb[1], wrgb[2]);
                    if (film_w) {
                        const Sp8 fr = film_of(hero, hp, dir, nr, mat);
                        for (int k = 0; k < 8; ++k) w.v[k] *= fr.v[k];
                    }
                    Sp8 dw = sp_splat(1.f);
                    if (collapse >= 0)
                        for (int k = 0; k < 8; ++k) dw.v[k] = k == collapse ? 8.f : 0.f;  // spec_collapse_mask
                    if (soft_w) {
                        // dispersion_angles2 + dispersion_weights (template.wgsl:3521-3548)
                        float lam[8];
                        lambdas_of(hero, lam);
                        const float ior_d = fmaxf(mat.ior, 1.f);
                        const float kk = -0.5f / fmaxf(S.disp_soft * S.disp_soft, 1e-12f);
                        float total = 0.f;
                        for (int k = 0; k < 8; ++k) {
                            const float ior_k = fmaxf(cauchy_ior(ior_d, mat.abbe, lam[k] * 0.001f), 1.f);
                            bool tk;
                            const V3<float> dk = refract_safe(dir, nr, in_solid ? ior_k : 1.f / ior_k, &tk);
                            const float ang = 2.f * (1.f - fclamp(dot3(nd, dk), -1.f, 1.f));
                            dw.v[k] = expf(ang * kk);
                            total += dw.v[k];
                        }
                        const float sc8 = 8.f / fmaxf(total, 1e-6f);
                        for (int k = 0; k < 8; ++k) dw.v[k] *= sc8;
                    }
                    for (int k = 0; k < 8; ++k) thr.v[k] *= dw.v[k] * w.v[k] * p_cont;
                    bool solid = in_solid;
                    int inner = a.P.inner[p];
                    if (kind == 2) {
                        solid = !in_solid;
                        if (solid) inner = min(max((int)mat.inner_max_steps, 4), 512);  // MIN/MAX_INNER_STEPS
                    }
                    const float bias = fmaxf(0.003f, 2.f * eps);
                    cont = true;
                    if (depth >= 4) {
                        const float survive = fclamp(sp_max(thr), 0.05f, 1.f);
                        if (hash12w(add_rn(sx, 59.3f), add_rn(sy, 23.9f)) > survive) cont = false;
                        else for (int k = 0; k < 8; ++k) thr.v[k] *= 1.f / survive;
                    }
                    if (cont && depth + 1 < S.max_depth) {
                        a.P.px[p] = hp.x + nd.x * bias; a.P.py[p] = hp.y + nd.y * bias; a.P.pz[p] = hp.z + nd.z * bias;
                        a.P.dx[p] = nd.x; a.P.dy[p] = nd.y; a.P.dz[p] = nd.z;
                        a.P.depth[p] = (uint8_t)(depth + 1);
                        // a mirror reflection at depth 0 (template.wgsl:4117-4118)
                        const bool mirror = S.cache_on && depth == 0 && !in_solid && kind == 1 && rough * rough <= 1e-4f;
                        a.P.flags[p] = (uint8_t)(((reflected || kind == 1) ? 1u : 0u) | (solid ? 4u : 0u) | (mirror ? 16u : 0u) |
                                                 (lens ? 32u : 0u));
                        a.P.inner[p] = (uint16_t)inner;
                    } else {
                        cont = false;
                    }
                    for (int k = 0; k < 8; ++k) a.P.thr[8 * p + k] = thr.v[k];
                }
            }
            for (int k = 0; k < 8; ++k) a.P.rad[8 * p + k] = rad.v[k];
        }
        uint32_t slot = 0;
        append(a.P.ctr + C_LIVE_OUT, cont, &slot);
        if (cont) qout[slot] = p;
    }
    for (int off = 32; off > 0; off >>= 1) de += __shfl_down(de, off);
    if (threadIdx.x == 0 && de) atomicAdd(de_ctr(a), de);
}

// --- indirect shading: shade_local at the indirect hit, no shadows, ambient 0.15
// (indirect_light_traced, template.wgsl:3359-3377)
template <bool EXT>
__global__ void __launch_bounds__(64) k_mc_indirect_shade(McArgs a0, const DevScene* __restrict__ scene, const McScene* __restrict__ mcs) {
    McArgs a = restricted(a0, scene, mcs);
    const uint32_t n = a.P.ctr[C_IND];
    const DevJob& J = *a.job;
    const DevScene& sc = *a.scene;
    const McScene& S = *a.S;
    unsigned long long de = 0;
    for (uint32_t i = blockIdx.x * 64 + threadIdx.x; i < n; i += gridDim.x * 64) {
        const uint32_t p = a.P.ind_path[i];
        const Basis B = basis_of(a.P.hero[p]);
        const V3<float> dir = mk3(a.P.idx[i], a.P.idy[i], a.P.idz[i]);
        Sp8 L;
        if (!a.P.ihit[i]) {
            const V3<float> sky = sky_color(S, dir);
            L = from_rgb(B, sky.x, sky.y, sky.z);
        } else {
            const float t = a.P.it[i];
            const V3<float> bp = mk3(a.P.iox[i] + dir.x * t, a.P.ioy[i] + dir.y * t, a.P.ioz[i] + dir.z * t);
            const V3<float> bn = normal_tetra<EXT>(sc, bp, J.eps_c * t + J.eps_f);
            const Mat bm = part_mask(hit_material<EXT>(sc, S, bp), S);
            de += 5;
            L = shade_lights(a, B, bp, dir, bn, bm, film_of(a.P.hero[p], bp, dir, bn, bm), sp_splat(0.15f), false,
                             sp_splat(0.f), p, 0.f, 0.f);
        }
        for (int k = 0; k < 8; ++k) atomicAdd(&a.P.rad[8 * p + k], a.P.iw[8 * i + k] * L.v[k]);
    }
    for (int off = 32; off > 0; off >>= 1) de += __shfl_down(de, off);
    if (threadIdx.x == 0 && de) atomicAdd(de_ctr(a), de);
}

// --- shadow: visibility_march (template.wgsl:2282-2342) with shadow_factor's parameters,
// refilled. An opaque hit blocks the light. With TRANSP (scene_has_transparency), a hit on
// a surface whose material has transparency a >= 0.01 tints the ray by
// from_rgb(mix(1, colour, a)) * a in the path's spectral basis and steps through the solid
// (steps of max(|DE|, floor), floor 0.004 growing by 1.15) until DE > 0.001 t + 0.0005.
template <bool EXT, bool TRANSP>
__global__ void __launch_bounds__(64) k_mc_shadow(McArgs a0, const DevScene* __restrict__ scene, const McScene* __restrict__ mcs) {
    McArgs a = restricted(a0, scene, mcs);
    const DevScene& sc = *a.scene;
    const DevJob& J = *a.job;
    const bool acc = J.accel != nullptr;
    const AccelView av = accel_view(J);
    AccelCache ac;  // one grid for every shadow ray of the call
    accel_cache_reset(ac);
    const uint32_t n = a.P.ctr[C_SHADOW];
    const int max_steps = a.S->shadow_steps;
    uint32_t cnext = 0, cend = 0;
    bool work_left = true, active = false, first_claim = true;
    uint32_t item = 0;
    V3<float> so = mk3(0.f, 0.f, 0.f), ld = mk3(0.f, 0.f, 1.f);
    float t = 0.f, res = 1.f, ldist = 0.f, k = 0.f;
    int step = 0;
    bool ins = false;          // TRANSP: inside a transparent occluder
    float flo = 0.004f;        // SHADOW_STEP_FLOOR, grown by SHADOW_FLOOR_GROWTH inside
    Sp8 tint = sp_splat(1.f);  // TRANSP: accumulated transmission
    unsigned long long de = 0;
    for (;;) {
        if (work_left) {
            uint64_t need = __builtin_amdgcn_ballot_w64(!active);
            if (__popcll(need) >= 8 || need == ~0ull) {
                while (need != 0ull) {
                    if (cnext >= cend) {
                        const uint32_t base = claim_chunk(a.P.ctr + C_HEAD_SH, n, a.chunk, first_claim);
                        first_claim = false;
                        if (base >= n) { work_left = false; break; }
                        cnext = base;
                        cend = (uint32_t)__builtin_amdgcn_readfirstlane((int)min(base + a.chunk, n));
                    }
                    const uint32_t avail = cend - cnext;
                    const unsigned lane = __lane_id();
                    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
                    const uint32_t rank = (uint32_t)__popcll(need & lt);
                    const bool mine = ((need >> lane) & 1ull) && rank < avail;
                    if (mine) {
                        item = cnext + rank;
                        so = mk3(a.P.sox[item], a.P.soy[item], a.P.soz[item]);
                        ld = mk3(a.P.sdx[item], a.P.sdy[item], a.P.sdz[item]);
                        ldist = a.P.sdist[item];
                        k = a.P.sk[item];
                        t = 0.004f; res = 1.f; step = 0;
                        if (TRANSP) { ins = false; flo = 0.004f; tint = sp_splat(1.f); }
                        if (t < ldist) active = true;
                        else {
                            const uint32_t p = a.P.shadow_path[item];
                            for (int q = 0; q < 8; ++q) atomicAdd(&a.P.rad[8 * p + q], a.P.sw[8 * item + q]);
                        }
                    }
                    cnext += min((uint32_t)__popcll(need), avail);
                    need = __builtin_amdgcn_ballot_w64(((need >> lane) & 1ull) && !mine);
                }
            }
        }
        const uint64_t act = __builtin_amdgcn_ballot_w64(active);
        if (act == 0ull) {
            if (!work_left) break;
            continue;
        }
        bool done = false, skipped = false;
        if (active) {
            const V3<float> ps = mk3(so.x + ld.x * t, so.y + ld.y * t, so.z + ld.z * t);
            step++;
            float skip = 0.f;
            if (acc && !(TRANSP && ins)) {
                skip = accel_skip_cached(av, ac, ps);
                skipped = shadow_skip_applies(skip, k, res, t);
            }
            // a skip takes the step without a DE (visibility_march, template.wgsl:2297-2303)
            const float d = skipped ? 0.f : scene_de<false, EXT>(sc, ps, nullptr);
            if (skipped) {
                t += skip;
            } else if (TRANSP && ins) {
                if (d > 0.001f * t + 0.0005f) ins = false;  // SHADOW_EXIT_SLOPE, _BIAS
                t += fmaxf(fabsf(d), flo);
                flo *= 1.15f;
            } else if (d < 0.0005f) {  // SHADOW_HIT_EPS
                bool blocked = true;
                if (TRANSP) {
                    // the occluder's material, unmasked (hit_material(p).mat)
                    const Mat om = hit_material<EXT>(sc, *a.S, ps);
                    const float al = fclamp(om.transparency, 0.f, 1.f);
                    if (al >= 0.01f) {  // SHADOW_OPAQUE_ALPHA
                        const Basis B = basis_of(a.P.hero[a.P.shadow_path[item]]);
                        const Sp8 tc = from_rgb(B, 1.f + (om.color[0] - 1.f) * al, 1.f + (om.color[1] - 1.f) * al,
                                                1.f + (om.color[2] - 1.f) * al);
                        for (int q = 0; q < 8; ++q) tint.v[q] *= tc.v[q] * al;
                        if (sp_max(tint) >= 0.004f) {  // SHADOW_TINT_CUT
                            blocked = false;
                            ins = true;
                            t += flo;
                        }
                    }
                }
                if (blocked) { res = 0.f; done = true; }
            } else {
                res = fminf(res, k * d / t);
                t += fmaxf(d, 0.001f);
            }
            if (!done && (step >= max_steps || t >= ldist)) done = true;
            if (done) {
                const float vis = fclamp(res, 0.f, 1.f);
                if (vis > 0.f) {
                    const uint32_t p = a.P.shadow_path[item];
                    for (int q = 0; q < 8; ++q)
                        atomicAdd(&a.P.rad[8 * p + q], a.P.sw[8 * item + q] * vis * (TRANSP ? tint.v[q] : 1.f));
                }
                active = false;
            }
        }
        de += (unsigned long long)__popcll(act & ~__builtin_amdgcn_ballot_w64(skipped));
    }
    if (__lane_id() == 0 && de) atomicAdd(de_ctr(a), de);
}

// One sample's colour of a pixel: the gamut-fitted RGB of its path, or with depth of field
// the mean over its lens paths (fs_wf_resolve, template.wgsl:5373-5394).
NDD void sample_rgb(const McArgs& a, uint32_t s, uint32_t pix, float rgb[3]) {
    const uint32_t lens = pixel_lens_samples(a, pix);
    const uint32_t p0 = (s * a.npix + pix) * (uint32_t)a.S->slots;
    rgb[0] = rgb[1] = rgb[2] = 0.f;
    for (uint32_t j = 0; j < lens; ++j) {
        const uint32_t p = p0 + j;
        Sp8 r;
        for (int k = 0; k < 8; ++k) r.v[k] = a.P.rad[8 * p + k];
        float c[3];
        to_rgb_fit(a.P.hero[p], r, c);
        rgb[0] += c[0]; rgb[1] += c[1]; rgb[2] += c[2];
    }
    if (lens > 1u) {
        const float inv = 1.f / (float)lens;
        rgb[0] *= inv; rgb[1] *= inv; rgb[2] *= inv;
    }
}

// --- resolve: per pixel, the average over this chunk's samples is added into out_rgb
// (scaled by 1 / spp_total)
__global__ void k_mc_resolve(McArgs a) {
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= a.npix) return;
    float acc[3] = {0.f, 0.f, 0.f};
    const uint32_t cs = a.ch->spp_chunk;
    for (uint32_t s = 0; s < cs; ++s) {
        float rgb[3];
        sample_rgb(a, s, pix, rgb);
        acc[0] += rgb[0]; acc[1] += rgb[1]; acc[2] += rgb[2];
    }
    const float inv = 1.f / (float)a.ch->spp_total;
    for (int k = 0; k < 3; ++k) a.out_rgb[3 * pix + k] += acc[k] * inv;
}

// --- resolve with adaptive sampling (adapt_update, template.wgsl:4749-4760; the sample skip
// of cs_wf_generate, :5134-5146): per pixel, in sample order, a sample after convergence is
// dropped; the others are added to the pixel's colour sum and luminance statistics, and the
// output is the mean of the samples taken. The state persists across chunks and calls.
__global__ void k_mc_resolve_adapt(McArgs a) {
    const uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= a.npix) return;
    const McScene& S = *a.S;
    const DevJob& J = *a.job;
    float* st = a.adapt + 8 * (size_t)pix;
    float s4[4] = {st[0], st[1], st[2], st[3]};
    float sum[3] = {st[4], st[5], st[6]};
    float taken = st[7];
    const bool on = (float)(J.rect_x + (int)(pix % (uint32_t)J.rect_w)) < S.adapt_row_width;
    const float w[3] = {0.2126f, 0.7152f, 0.0722f};
    for (uint32_t s = 0; s < a.ch->spp_chunk; ++s) {
        const float ms = add_rn(a.ch->mc_sample0, (float)(a.ch->sample_base + s));
        if (on && ms > 0.5f && s4[3] > 0.5f) continue;
        float rgb[3];
        sample_rgb(a, s, pix, rgb);
        for (int k = 0; k < 3; ++k) sum[k] += rgb[k];
        taken += 1.f;
        if (on) {
            float l = 0.f;
            for (int k = 0; k < 3; ++k) l += rgb[k] / (rgb[k] + 1.f) * w[k];
            s4[0] += l; s4[1] += l * l; s4[2] += 1.f;
            if (s4[2] >= fmaxf(S.adapt_min_samples, 2.f)) {
                const float mean = s4[0] / s4[2];
                const float var = fmaxf(s4[1] / s4[2] - mean * mean, 0.f);
                s4[3] = sqrtf(var / s4[2]) <= S.adapt_threshold * (mean + 0.05f) ? 1.f : 0.f;
            }
        }
    }
    for (int k = 0; k < 4; ++k) st[k] = s4[k];
    for (int k = 0; k < 3; ++k) st[4 + k] = sum[k];
    st[7] = taken;
    const float inv = 1.f / fmaxf(taken, 1.f);
    for (int k = 0; k < 3; ++k) a.out_rgb[3 * pix + k] = sum[k] * inv;
}

// counters between passes (no host round trip)
__global__ void k_mc_counters(uint32_t* ctr, int op) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (op == 3) {  // new call: totals too
        for (int i = 0; i < 16; ++i) ctr[i] = 0;
    } else if (op == 0) {  // new chunk: all queues empty (the totals C_SEG.. are kept)
        for (int i = 0; i < C_SEG; ++i) ctr[i] = 0;
    } else if (op == 1) {  // before a bounce: clear the out queue and the side queues
        ctr[C_LIVE_OUT] = 0; ctr[C_SHADOW] = 0; ctr[C_IND] = 0;
        ctr[C_HEAD_EXT] = 0; ctr[C_HEAD_SH] = 0; ctr[C_HEAD_IND] = 0;
    } else if (op == 2) {  // after a bounce: the out queue becomes the in queue
        ctr[C_SEG] += ctr[C_LIVE_IN];
        ctr[C_SHTOT] += ctr[C_SHADOW];
        ctr[C_LIVE_IN] = ctr[C_LIVE_OUT];
    }
}
// the live queue of a chunk: every path, or (adaptive sampling) none yet, generate appends
__global__ void k_mc_set_live(uint32_t* ctr, const McChunk* ch, uint32_t npix, int adapt) {
    if (threadIdx.x == 0 && blockIdx.x == 0) ctr[C_LIVE_IN] = adapt ? 0u : npix * ch->spp_chunk;
}
// Copies chunk record k into the current-chunk slot that the captured passes read.
__global__ void k_mc_select_chunk(McChunk* recs, uint32_t k) {
    if (threadIdx.x == 0 && blockIdx.x == 0) recs[MC_MAX_CHUNKS] = recs[k];
}
__global__ void k_mc_zero(float* p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = 0.f;
}

}  // namespace nd

// ----------------------------------------------------------------------------- host
static void pack_mc_scene(const nd_uniforms& u, nd::McScene& S) {
    std::memset(&S, 0, sizeof(S));
    const uint32_t parts = (uint32_t)u.debug_parts;
    for (int i = 0; i < ND_MAX_INSTANCES; ++i) {
        const nd_fractal_instance& fi = u.instances[i];
        nd::McInst& I = S.inst[i];
        I.color_count = (int)fi.color_count;
        I.trap_repeat = (int)(fi.trap_repeat + 0.5f);
        I.trap_span = fi.trap_span;
        I.trap_offset = fi.trap_offset;
        for (int s = 0; s < 16; ++s) {
            const nd_color_stop& c = fi.colors[s];
            nd::McStop& d = I.stops[s];
            for (int k = 0; k < 3; ++k) d.color[k] = c.color[k];
            d.position = c.position; d.glossiness = c.glossiness; d.transparency = c.transparency;
            d.reflectiveness = c.reflectiveness; d.ior = c.ior; d.roughness = c.roughness;
            d.abbe = c.abbe; d.inner_max_steps = c.inner_max_steps;
            d.film_thickness = c.film_thickness; d.film_ior = c.film_ior; d.film_strength = c.film_strength;
            d.film_angle_scale = c.film_angle_scale; d.film_perturb = c.film_perturb;
            d.film_perturb_scale = c.film_perturb_scale;
            d.subsurface = c.subsurface; d.subsurface_radius = c.subsurface_radius;
        }
    }
    S.light_count = (parts & ND_PART_DIRECT_LIGHTS) ? 0 : std::min((int)u.light_count, ND_MAX_LIGHTS);
    for (int i = 0; i < S.light_count; ++i) {
        const nd_light& L = u.lights[i];
        nd::McLight& d = S.lights[i];
        const int lt = (int)(L.light_type + 0.5f);
        d.type = lt == 1 ? 1 : lt == 2 ? 2 : 0;
        d.cast = L.cast_shadows > 0.5f;
        for (int k = 0; k < 3; ++k) d.color[k] = L.color[k] * L.brightness;
        if (d.type == 1) {
            double n = std::sqrt((double)L.position_or_direction[0] * L.position_or_direction[0] +
                                 (double)L.position_or_direction[1] * L.position_or_direction[1] +
                                 (double)L.position_or_direction[2] * L.position_or_direction[2]);
            for (int k = 0; k < 3; ++k) d.v[k] = (float)(L.position_or_direction[k] / n);
        } else {
            for (int k = 0; k < 3; ++k) d.v[k] = L.position_or_direction[k];
        }
        d.k = L.shadow_softness + (512.0f - L.shadow_softness) * L.hard_shadows;
        // ray light: axis = ray_direction / max(|ray_direction|, EPSILON_FINE), spread <= 1.5 rad
        const double al = std::sqrt((double)L.ray_direction[0] * L.ray_direction[0] +
                                    (double)L.ray_direction[1] * L.ray_direction[1] +
                                    (double)L.ray_direction[2] * L.ray_direction[2]);
        for (int k = 0; k < 3; ++k) d.axis[k] = (float)(L.ray_direction[k] / std::max(al, 1e-6));
        d.waist = L.waist;
        d.tan_spread = std::tan(std::min(std::max(L.spread, 0.0f), 1.5f));
    }
    S.shadows_on = !(parts & ND_PART_SHADOWS);
    S.specular_on = !(parts & ND_PART_SPECULAR);
    S.ao_on = !(parts & ND_PART_AMBIENT_OCCLUSION);
    S.reflections_on = !(parts & ND_PART_REFLECTIONS);
    S.indirect_on = u.mc_enabled > 0.5f && !(parts & ND_PART_MC_INDIRECT);
    S.sky_on = !(parts & ND_PART_SKY);
    S.strips_on = !(parts & ND_PART_COLOR_STRIPS);
    S.refraction_on = !(parts & ND_PART_REFRACTION);
    S.dispersion_on = !(parts & ND_PART_DISPERSION);
    S.iridescence_on = !(parts & ND_PART_IRIDESCENCE);
    S.subsurface_on = !(parts & ND_PART_SUBSURFACE);
    S.simple_view = (parts & ND_PART_GEOMETRY_ONLY) != 0;
    S.sss = 0;
    if (S.subsurface_on)
        for (int i = 0; i < ND_MAX_INSTANCES; ++i)
            for (int c = 0; c < std::min((int)u.instances[i].color_count, 16); ++c)
                if (u.instances[i].colors[c].subsurface > 0.001f) S.sss = 1;  // MATERIAL_EPSILON
    S.sss_steps = std::min((int)u.max_steps, 128);
    S.transp = u.has_transparency > 0.5f && S.refraction_on;
    S.transmissive = 0;
    if (S.refraction_on)
        for (int i = 0; i < ND_MAX_INSTANCES; ++i)
            for (int c = 0; c < std::min((int)u.instances[i].color_count, 16); ++c)
                if (u.instances[i].colors[c].transparency > 0.f) S.transmissive = 1;
    S.light_pick = u.mc_enabled > 0.5f && u.mc_light_pick > 0.5f;
    S.cache_on = u.primary_cache_key > 0.5f;
    S.adapt_on = u.adapt_enabled > 0.5f && u.mc_enabled > 0.5f;
    S.adapt_row_width = u.adapt_row_width;
    S.adapt_min_samples = u.adapt_min_samples;
    S.adapt_threshold = u.adapt_threshold;
    S.disp_soft = (0.5f * u.photon_dispersion_soft) * 0.017453292519943295f;
    S.max_depth = std::min(std::max((int)u.max_reflection_bounces + 8, 1), 16);
    S.shadow_steps = std::min((int)u.max_steps, 40);
    for (int k = 0; k < 3; ++k) {
        S.sky_zenith[k] = u.sky.zenith[k]; S.sky_horizon[k] = u.sky.horizon[k]; S.sky_ground[k] = u.sky.ground[k];
        S.sky_sun_color[k] = u.sky.sun_color[k];
    }
    double sn = std::sqrt((double)u.sky.sun_dir[0] * u.sky.sun_dir[0] + (double)u.sky.sun_dir[1] * u.sky.sun_dir[1] +
                          (double)u.sky.sun_dir[2] * u.sky.sun_dir[2]);
    for (int k = 0; k < 3; ++k) S.sky_sun_dir[k] = sn > 0 ? (float)(u.sky.sun_dir[k] / sn) : 0.f;
    S.sky_yaw_c = std::cos(u.sky.yaw);
    S.sky_yaw_s = std::sin(u.sky.yaw);
    S.sky_inv_falloff = 1.0f / std::max(u.sky.falloff, 0.01f);
    S.sky_sun = u.sky.sun_cos > -1.0f;
    S.sky_sun_cos = u.sky.sun_cos;
    S.sky_sun_intensity = u.sky.sun_intensity;
    S.sky_intensity = u.sky.intensity;
    S.max_dist = u.max_dist;
    // depth of field (cs_wf_generate, template.wgsl:5150-5182): the wavefront's inner slots
    // per pixel (wavefrontInner, fractal_renderer.zig:4335-4337) and mc_inner_count(32)
    S.dof_on = u.dof_enabled > 0.5f && u.aperture > 0.0001f && !(parts & ND_PART_DEPTH_OF_FIELD);
    const int inner = std::min(std::max((int)std::max(u.mc_inner_samples, 1.0f), 1), 32);
    const int n_dof = u.mc_enabled > 0.5f ? std::min(std::max((int)u.mc_inner_samples, 1), 32) : 32;
    S.slots = S.dof_on ? inner : 1;
    S.dof_n = std::min(n_dof, inner);
    S.aperture = u.aperture;
    S.focus_distance = u.focus_distance;
    S.focus_range = u.focus_range;
}

namespace nd {
// ----------------------------------------------------------------------------- G-buffer colour
// undo_tonemap (template.wgsl:4139-4142): the inverse of c / (1 + c), clamped
NDD V3<float> undo_tonemap(V3<float> c) {
    const float k = 1.f / 255.f;  // 1 / UNORM8_MAX
    c = mk3(fclamp(c.x, 0.f, 1.f), fclamp(c.y, 0.f, 1.f), fclamp(c.z, 0.f, 1.f));
    return mk3(c.x / fmaxf(1.f - c.x, k), c.y / fmaxf(1.f - c.y, k), c.z / fmaxf(1.f - c.z, k));
}
// fs_simple's colour attachment and fs_aov's albedo for one pixel of every job, from the
// geometry passes' hit and depth (template.wgsl:4417-4538, 4853-4896). One thread per ray
// slot of the batch (64 per 8x8 tile of each job's rect).
template <bool EXT>
__global__ void __launch_bounds__(256) k_gbuffer_color(const DevJob* __restrict__ jobs, const McScene* __restrict__ scenes,
                                                      const uint32_t* __restrict__ job_base, int n_jobs, uint32_t total) {
    const uint32_t slot = blockIdx.x * 256 + threadIdx.x;
    if (slot >= total) return;
    const int job = find_job(job_base, n_jobs, slot);
    const DevJob& J = jobs[job];
    int x, y;
    if (!slot_pixel(J, slot - J.ray_base, &x, &y)) return;
    const McScene& S = scenes[job];
    const size_t o = (size_t)y * J.rect_w + x;
    const V3<float> dir = primary_dir(J, J.rect_x + x, J.rect_y + y);
    const V3<float> bg = mk3(0.10f, 0.11f, 0.13f);  // SIMPLE_BACKGROUND
    V3<float> color, albedo;
    if (!J.out_hit[o]) {
        const V3<float> sky = sky_color(S, dir);
        color = S.simple_view ? undo_tonemap(bg) : sky;
        albedo = mk3(fclamp(sky.x, 0.f, 1.f), fclamp(sky.y, 0.f, 1.f), fclamp(sky.z, 0.f, 1.f));
    } else {
        const DevScene& sc = J.scene;
        const float t = J.out_depth[o];
        const V3<float> hp = mk3(J.cam[0] + dir.x * t, J.cam[1] + dir.y * t, J.cam[2] + dir.z * t);
        const V3<float> n = normal_tetra<EXT>(sc, hp, J.eps_c * t + J.eps_f);  // estimate_normal
        const Mat m = part_mask(hit_material<EXT>(sc, S, hp), S);
        if (S.simple_view) {
            // simple_geometry: a key light from the camera (SIMPLE_KEY_UP 0.45, _LEFT 0.35),
            // ambient 0.18, albedo 0.72/0.72/0.73, depth cue 0.55 toward the background
            const V3<float> kd = vnorm3(mk3(-dir.x + J.up[0] * 0.45f - J.right[0] * 0.35f,
                                            -dir.y + J.up[1] * 0.45f - J.right[1] * 0.35f,
                                            -dir.z + J.up[2] * 0.45f - J.right[2] * 0.35f));
            const float shade = 0.18f + (1.f - 0.18f) * fmaxf(dot3(n, kd), 0.f);
            const float cue = fclamp(t / fmaxf(J.max_dist, 1e-4f), 0.f, 1.f) * 0.55f;
            color = undo_tonemap(mk3(0.72f * shade * (1.f - cue) + bg.x * cue, 0.72f * shade * (1.f - cue) + bg.y * cue,
                                     0.73f * shade * (1.f - cue) + bg.z * cue));
        } else {
            // shade_lite (template.wgsl:4424-4444): ambient 0.15 x AO, every light without a
            // shadow, Blinn-Phong with the half vector of l and -d, all times 1 - Fresnel
            float ao = 1.f;
            if (S.ao_on) {
                const float radius = fclamp(t * 0.04f, 0.003f, 0.25f);
                float occ = 0.f, w = 1.f;
                for (int k = 0; k < 5; ++k) {
                    const float step = radius * (0.2f + 0.2f * (float)k);
                    const float dd = scene_de<false, EXT>(sc, mk3(hp.x + n.x * step, hp.y + n.y * step, hp.z + n.z * step), nullptr);
                    occ += fmaxf(step - dd, 0.f) * w;
                    w *= 0.6f;
                }
                ao = fclamp(1.f - 2.f * occ / radius, 0.f, 1.f);
            }
            const float ior = fmaxf(m.ior, 1.f);
            const float r0r = (1.f - ior) / (1.f + ior), r0 = r0r * r0r;
            const float cth = fclamp(1.f - fabsf(dot3(dir, n)), 0.f, 1.f), c2 = cth * cth;
            const float fres = r0 + (1.f - r0) * c2 * c2 * cth;
            const float power = 4.f + 124.f * fclamp(m.glossiness, 0.f, 1.f);
            V3<float> dif = mk3(0.15f * ao, 0.15f * ao, 0.15f * ao), spc = mk3(0.f, 0.f, 0.f);
            for (int i = 0; i < S.light_count; ++i) {
                V3<float> ld, lc;
                float ldist;
                light_of(S.lights[i], S.max_dist, hp, &ld, &lc, &ldist);
                const float ndotl = fmaxf(dot3(n, ld), 0.f);
                if (ndotl <= 0.f) continue;
                dif = mk3(dif.x + lc.x * ndotl, dif.y + lc.y * ndotl, dif.z + lc.z * ndotl);
                if (S.specular_on) {
                    const V3<float> h = vnorm3(mk3(ld.x - dir.x, ld.y - dir.y, ld.z - dir.z));
                    const float sp = powf(fmaxf(dot3(n, h), 0.f), power) * m.glossiness;
                    spc = mk3(spc.x + lc.x * sp, spc.y + lc.y * sp, spc.z + lc.z * sp);
                }
            }
            const float f1 = 1.f - fres;
            color = mk3(f1 * (m.color[0] * dif.x + spc.x), f1 * (m.color[1] * dif.y + spc.y), f1 * (m.color[2] * dif.z + spc.z));
        }
        // aov_albedo: the colour mixed toward white by max(reflectiveness, transparency)
        const float sp = fclamp(fmaxf(m.reflectiveness, m.transparency), 0.f, 1.f);
        albedo = mk3(fclamp(m.color[0] + (1.f - m.color[0]) * sp, 0.f, 1.f), fclamp(m.color[1] + (1.f - m.color[1]) * sp, 0.f, 1.f),
                     fclamp(m.color[2] + (1.f - m.color[2]) * sp, 0.f, 1.f));
    }
    if (J.out_color) { J.out_color[3 * o] = color.x; J.out_color[3 * o + 1] = color.y; J.out_color[3 * o + 2] = color.z; }
    if (J.out_albedo) { J.out_albedo[3 * o] = albedo.x; J.out_albedo[3 * o + 1] = albedo.y; J.out_albedo[3 * o + 2] = albedo.z; }
}

// ----------------------------------------------------------------------------- radiance adjoint
// d(direct-lighting colour)/d(theta), theta = (Mandelbulb power, Mandelbox scale, Mandelbox min
// radius, camera x, y, z), for the hits of a geometry batch. The hit is first polished to the
// surface DE(o + t d) = eps(t) by Newton steps from the march's hit (which its bisection leaves
// up to eps(t) past it): the point the implicit-function tangent describes, and the one the
// march converges to as its bisection window shrinks. The colour there is the path trace's at
// depth 0 with Monte Carlo off and the reflection continuation cut: the scene's constant
// material, Schlick's local share, ambient 0.15 x AO and every light with its soft shadow,
// through the hero wavelength's spectrum to RGB and gamut_fit (template.wgsl:3819-4137).
// Tangent slots: 0..2 camera (and space, for the hit depth's implicit-function tangent),
// 3..5 the formula parameters the batch's jobs carry as adjoint slots.
using D6 = DualN<6>;
NDD V3<D6> d3(const V3<float>& v) { return mk3(dconst<6>(v.x), dconst<6>(v.y), dconst<6>(v.z)); }
NDD V3<D6> dadd(const V3<D6>& a, const V3<D6>& b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
NDD V3<D6> dscale(const V3<D6>& a, const D6& s) { return mk3(a.x * s, a.y * s, a.z * s); }
NDD V3<D6> dnorm(const V3<D6>& g) {
    const D6 l = ssqrt(dot3(g, g));
    return mk3(g.x / l, g.y / l, g.z / l);
}
NDD D6 dclamp(const D6& x, float lo, float hi) { return sclamp(x, dconst<6>(lo), dconst<6>(hi)); }
NDD D6 dmaxf(const D6& x, float b) { return smax_(x, dconst<6>(b)); }

// light_of with dual positions: direction, colour scale a (colour = L.color * a) and distance
NDD void light_of_d(const McLight& L, float max_dist, const V3<D6>& p, V3<D6>* dir, D6* a, D6* dist) {
    if (L.type == 1) {
        *dir = d3(mk3(L.v[0], L.v[1], L.v[2]));
        *a = dconst<6>(1.f);
        *dist = dconst<6>(max_dist * 2.f);
        return;
    }
    const V3<D6> to = mk3(dconst<6>(L.v[0]) - p.x, dconst<6>(L.v[1]) - p.y, dconst<6>(L.v[2]) - p.z);
    const D6 d = ssqrt(dot3(to, to));
    const D6 inv = dconst<6>(1.f) / dmaxf(d, 1e-4f);
    *dir = dscale(to, inv);
    *dist = d;
    if (L.type == 0) {
        *a = dconst<6>(1.f) / dmaxf(d * d, 0.01f);
    } else {
        const V3<D6> ax = d3(mk3(L.axis[0], L.axis[1], L.axis[2]));
        const D6 t0 = -dot3(to, ax);
        const V3<D6> q = mk3(-to.x - ax.x * t0, -to.y - ax.y * t0, -to.z - ax.z * t0);
        const D6 w = dmaxf(t0, 0.f) * L.tan_spread + L.waist;
        const D6 w2 = w * w;
        const D6 e = -dot3(q, q) / w2;
        const D6 profile = dchain(e, expf(e.v), expf(e.v));
        *a = t0.v > 0.f ? profile / w2 : dconst<6>(0.f);
    }
}
// the opaque soft shadow of k_mc_shadow with dual origin, direction and distance
template <bool EXT>
NDD D6 vis_d(const DevScene& sc, const V3<D6>& o, const V3<D6>& ld, const D6& ldist, float k, int steps) {
    D6 t = dconst<6>(0.004f), res = dconst<6>(1.f);
    if (!(t.v < ldist.v)) return res;
    for (int step = 1; step <= steps; ++step) {
        const V3<D6> p = dadd(o, dscale(ld, t));
        const D6 d = scene_de_d<D6, EXT>(sc, p);
        if (d.v < 0.0005f) return dconst<6>(0.f);  // SHADOW_HIT_EPS
        res = smin_(res, d * k / t);
        t = t + dmaxf(d, 0.001f);
        if (!(t.v < ldist.v)) break;
    }
    return dclamp(res, 0.f, 1.f);
}
// the spectrum -> RGB weights of to_rgb_fit (before gamut_fit)
NDD void rgb_weights(float hero, float* wr, float* wg, float* wb) {
    float l[8];
    lambdas_of(hero, l);
    float sr = 0.f, sg = 0.f, sb = 0.f;
    for (int k = 0; k < 8; ++k) {
        float x = 1.056f * cie_lobe(l[k], 599.8f, 37.9f, 31.0f) + 0.362f * cie_lobe(l[k], 442.0f, 16.0f, 26.7f) -
                  0.065f * cie_lobe(l[k], 501.1f, 20.4f, 26.2f);
        float y = 0.821f * cie_lobe(l[k], 568.8f, 46.9f, 40.5f) + 0.286f * cie_lobe(l[k], 530.9f, 16.3f, 31.1f);
        float z = 1.217f * cie_lobe(l[k], 437.0f, 11.8f, 36.0f) + 0.681f * cie_lobe(l[k], 459.0f, 26.0f, 13.8f);
        wr[k] = 3.2404542f * x - 1.5371385f * y - 0.4985314f * z;
        wg[k] = -0.9692660f * x + 1.8760108f * y + 0.0415560f * z;
        wb[k] = 0.0556434f * x - 0.2040259f * y + 1.0572252f * z;
        sr += wr[k]; sg += wg[k]; sb += wb[k];
    }
    for (int k = 0; k < 8; ++k) { wr[k] /= sr; wg[k] /= sg; wb[k] /= sb; }
}
NDD V3<float> lin_rgb(const float* wr, const float* wg, const float* wb, const Sp8& s) {
    float c0 = 0.f, c1 = 0.f, c2 = 0.f;
    for (int k = 0; k < 8; ++k) { c0 += s.v[k] * wr[k]; c1 += s.v[k] * wg[k]; c2 += s.v[k] * wb[k]; }
    return mk3(c0, c1, c2);
}

template <bool EXT>
__global__ void __launch_bounds__(64) k_radiance_adjoint(NdRadianceLaunch a) {
    const uint32_t n_hits = a.ctr->hit_count;
    const int slot_of[6] = {3, 4, 5, 0, 1, 2};  // theta order -> tangent slot
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = blockIdx.x * 64 + threadIdx.x; i - threadIdx.x < n_hits; i += gridDim.x * 64) {
        if (i >= n_hits) continue;
        const uint32_t r = a.hb.ray[i];
        if (r == ~0u) continue;
        const int job = a.hb.job[i];
        const DevJob& J = a.jobs[job];
        const DevScene& sc = a.shared ? *a.shared : J.scene;
        const McScene& S = ((const McScene*)a.scenes)[job];
        int x, y;
        slot_pixel(J, r - J.ray_base, &x, &y);
        const size_t o = (size_t)y * J.rect_w + x;
        const int gx = J.rect_x + x, gy = J.rect_y + y;
        const V3<float> dir = primary_dir(J, gx, gy);
        // F(t, theta) = DE(o + t d; theta) - eps_c t - eps_f: Newton from the march's hit
        // (F <= 0 there, F' < 0 going in), kept inside [t / 2, t + eps(t)], then the hit depth's
        // tangent dt/dtheta = -dF/dtheta / F' from the last evaluation
        const float tm = a.hb.t[i];
        float t = tm;
        D6 de;
        float den = 0.f;
        for (int it = 0; it < 4; ++it) {
            const V3<float> q = mk3(J.cam[0] + dir.x * t, J.cam[1] + dir.y * t, J.cam[2] + dir.z * t);
            de = scene_de_d<D6, EXT>(sc, mk3(dvar<6>(q.x, 0), dvar<6>(q.y, 1), dvar<6>(q.z, 2)));
            den = de.d[0] * dir.x + de.d[1] * dir.y + de.d[2] * dir.z - J.eps_c;
            if (it == 3 || !(den < 0.f)) break;
            const float F = de.v - (J.eps_c * t + J.eps_f);
            t = fminf(fmaxf(t - F / den, 0.5f * tm), tm + J.eps_c * tm + J.eps_f);
        }
        if (den == 0.f) continue;
        const V3<float> hp = mk3(J.cam[0] + dir.x * t, J.cam[1] + dir.y * t, J.cam[2] + dir.z * t);
        D6 td = dconst<6>(t);
#pragma unroll
        for (int k = 0; k < 6; ++k) td.d[k] = -de.d[k] / den;
        const V3<D6> od = mk3(dvar<6>(J.cam[0], 0), dvar<6>(J.cam[1], 1), dvar<6>(J.cam[2], 2));
        const V3<D6> dd = d3(dir);
        const V3<D6> hpd = dadd(od, dscale(dd, td));
        // estimate_normal at e = eps(t)
        const D6 e = td * J.eps_c + J.eps_f;
        const float K[4][3] = {{1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {1, 1, 1}};
        V3<D6> g = d3(mk3(0.f, 0.f, 0.f));
        for (int q = 0; q < 4; ++q) {
            const V3<D6> pq = mk3(hpd.x + e * K[q][0], hpd.y + e * K[q][1], hpd.z + e * K[q][2]);
            const D6 dq = scene_de_d<D6, EXT>(sc, pq);
            g = mk3(g.x + dq * K[q][0], g.y + dq * K[q][1], g.z + dq * K[q][2]);
        }
        const V3<D6> nd = dnorm(g);
        const Mat m = part_mask(hit_material<EXT>(sc, S, hp), S);  // constant over the surface
        const V3<D6> nf = dot3(dd, nd).v > 0.f ? mk3(-nd.x, -nd.y, -nd.z) : nd;
        const float ior = fmaxf(m.ior, 1.f);
        const float r0r = (1.f - ior) / (1.f + ior), r0 = r0r * r0r;
        const D6 cth = dclamp(dconst<6>(1.f) - sabs(dot3(dd, nf)), 0.f, 1.f);
        const D6 c2 = cth * cth;
        const D6 f = c2 * c2 * cth * (1.f - r0) + r0;
        const float transparency = fclamp(m.transparency, 0.f, 1.f);
        const float refl = fclamp(m.reflectiveness, 0.f, 1.f);
        const D6 p_reflect = dclamp(f * (1.f - refl) + refl, 0.f, 1.f);
        const D6 share = (dconst<6>(1.f) - p_reflect) * (1.f - transparency);
        V3<D6> c = d3(mk3(0.f, 0.f, 0.f));
        if (share.v > 0.001f) {
            const float bx = (float)gx + 0.5f, by = (float)gy + 0.5f;  // mc_sample 0
            const float hero = hash12w(add_rn(bx, 19.31f), add_rn(by, 47.11f));
            const Basis B = basis_of(hero);
            float wr[8], wg[8], wb[8];
            rgb_weights(hero, wr, wg, wb);
            const Sp8 albedo = from_rgb(B, m.color[0], m.color[1], m.color[2]);
            D6 ao = dconst<6>(1.f);
            if (S.ao_on) {
                const D6 radius = dclamp(td * 0.04f, 0.003f, 0.25f);
                D6 occ = dconst<6>(0.f);
                float w = 1.f;
                for (int q = 0; q < 5; ++q) {
                    const D6 step = radius * (0.2f + 0.2f * (float)q);
                    const D6 dq = scene_de_d<D6, EXT>(sc, dadd(hpd, dscale(nd, step)));
                    occ = occ + dmaxf(step - dq, 0.f) * w;
                    w *= 0.6f;
                }
                ao = dclamp(dconst<6>(1.f) - occ * 2.f / radius, 0.f, 1.f);
            }
            const V3<float> ra = lin_rgb(wr, wg, wb, albedo);
            c = mk3(ao * (0.15f * ra.x), ao * (0.15f * ra.y), ao * (0.15f * ra.z));
            const float spec_power = 4.f + 124.f * fclamp(m.glossiness, 0.f, 1.f);
            for (int l = 0; l < S.light_count; ++l) {
                const McLight& L = S.lights[l];
                V3<D6> ld;
                D6 la, ldist;
                light_of_d(L, S.max_dist, hpd, &ld, &la, &ldist);
                if (fmaxf(L.color[0], fmaxf(L.color[1], L.color[2])) * la.v < 1e-5f) continue;
                const D6 ndotl = dmaxf(dot3(nd, ld), 0.f);
                if (!(ndotl.v > 0.f)) continue;
                D6 spec = dconst<6>(0.f);
                if (S.specular_on) {
                    const V3<D6> hd = dnorm(mk3(ld.x - dir.x, ld.y - dir.y, ld.z - dir.z));
                    const D6 nh = dmaxf(dot3(nd, hd), 0.f);
                    spec = spow(nh, dconst<6>(spec_power)) * m.glossiness;
                }
                D6 vis = dconst<6>(1.f);
                if (S.shadows_on && L.cast && S.shadow_steps > 0) {
                    const V3<D6> so = dadd(hpd, dscale(nd, dconst<6>(0.004f)));
                    vis = vis_d<EXT>(sc, so, ld, ldist, L.k, S.shadow_steps);
                    if (!(vis.v > 0.f)) continue;
                }
                const Sp8 ls = from_rgb(B, L.color[0], L.color[1], L.color[2]);
                Sp8 al;
                for (int k = 0; k < 8; ++k) al.v[k] = albedo.v[k] * ls.v[k];
                const V3<float> rd = lin_rgb(wr, wg, wb, al), rs = lin_rgb(wr, wg, wb, ls);
                const D6 w = la * vis;
                const D6 wd = w * ndotl, wsp = w * spec;
                c = mk3(c.x + wd * rd.x + wsp * rs.x, c.y + wd * rd.y + wsp * rs.y, c.z + wd * rd.z + wsp * rs.z);
            }
            c = dscale(c, share);
            // gamut_fit
            const D6 yl = c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f;
            if (!(yl.v > 0.f)) {
                c = d3(mk3(0.f, 0.f, 0.f));
            } else {
                const D6 mn = smin_(c.x, smin_(c.y, c.z));
                if (mn.v < 0.f) {
                    const D6 fit = yl / (yl - mn);
                    c = mk3(yl + (c.x - yl) * fit, yl + (c.y - yl) * fit, yl + (c.z - yl) * fit);
                }
            }
        }
        const uint64_t pix = a.pix_base[job] + o;
        const D6 cc[3] = {c.x, c.y, c.z};
        if (a.rgb)
            for (int q = 0; q < 3; ++q) a.rgb[pix * 3 + q] = cc[q].v;
        if (a.drgb)
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int k = 0; k < 6; ++k) a.drgb[pix * 18 + q * 6 + k] = cc[q].d[slot_of[k]];
        if (a.target && a.target[pix * 3] >= 0.f) {
            for (int q = 0; q < 3; ++q) {
                const double res = (double)cc[q].v - (double)a.target[pix * 3 + q];
                acc[0] += res * res;
#pragma unroll
                for (int k = 0; k < 6; ++k) acc[1 + k] += 2.0 * res * (double)cc[q].d[slot_of[k]];
            }
            acc[7] += 1.0;
        }
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        double v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
        if (threadIdx.x == 0 && v != 0.0) atomicAdd(&a.acc[k], v);
    }
}
}  // namespace nd

static_assert(sizeof(nd::McScene) <= ND_MC_SCENE_MAX_BYTES, "ND_MC_SCENE_MAX_BYTES");
size_t nd_mc_scene_bytes() { return sizeof(nd::McScene); }
void nd_mc_scene_pack(const nd_uniforms& u, void* out) { pack_mc_scene(u, *(nd::McScene*)out); }
hipError_t nd_launch_radiance_adjoint(const NdRadianceLaunch& L, uint32_t grid, bool ext, hipStream_t s) {
    if (ext) hipLaunchKernelGGL(nd::k_radiance_adjoint<true>, dim3(grid), dim3(64), 0, s, L);
    else hipLaunchKernelGGL(nd::k_radiance_adjoint<false>, dim3(grid), dim3(64), 0, s, L);
    return hipGetLastError();
}
uint32_t nd_resident_radiance(bool ext) {
    int b = 0;
    const hipError_t e = ext ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, nd::k_radiance_adjoint<true>, 64, 0)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, nd::k_radiance_adjoint<false>, 64, 0);
    return e == hipSuccess && b > 0 ? (uint32_t)b : 4u;
}
hipError_t nd_launch_gbuffer_color(const DevJob* jobs, const void* scenes, const uint32_t* job_base, int n_jobs,
                                   uint32_t total, bool ext, hipStream_t s) {
    const dim3 grid((total + 255) / 256), block(256);
    if (ext)
        hipLaunchKernelGGL(nd::k_gbuffer_color<true>, grid, block, 0, s, jobs, (const nd::McScene*)scenes, job_base, n_jobs, total);
    else
        hipLaunchKernelGGL(nd::k_gbuffer_color<false>, grid, block, 0, s, jobs, (const nd::McScene*)scenes, job_base, n_jobs, total);
    return hipGetLastError();
}

template <typename K>
static uint32_t mc_resident(K kernel) {
    int b = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, kernel, 64, 0) == hipSuccess && b > 0 ? (uint32_t)b : 8u;
}

void nd_resident_mc(uint32_t* r) {
    r[RES_MC_EXTEND] = mc_resident(nd::k_mc_extend<false, false>);
    r[RES_MC_SHADE] = mc_resident(nd::k_mc_shade<false, false>);
    r[RES_MC_SHADOW] = mc_resident(nd::k_mc_shadow<false, false>);
    r[RES_MC_IND_MARCH] = mc_resident(nd::k_mc_indirect_march<false>);
    r[RES_MC_IND_SHADE] = mc_resident(nd::k_mc_indirect_shade<false>);
    r[RES_MC_EXTEND_EXT] = mc_resident(nd::k_mc_extend<true, false>);
    r[RES_MC_SHADE_EXT] = mc_resident(nd::k_mc_shade<true, false>);
    r[RES_MC_SHADOW_EXT] = mc_resident(nd::k_mc_shadow<true, false>);
    r[RES_MC_IND_MARCH_EXT] = mc_resident(nd::k_mc_indirect_march<true>);
    r[RES_MC_IND_SHADE_EXT] = mc_resident(nd::k_mc_indirect_shade<true>);
}

extern "C" nd_status nd_path_trace(nd_ctx* c, const void* uniforms, size_t size, uint32_t spp, nd_rect rect,
                                   const nd_image_out* out, nd_trace_stats* stats) {
    if (!c) return ND_E_INVALID;
    if (!uniforms || size != sizeof(nd_uniforms)) return nd_fail(c, ND_E_INVALID, "uniforms_size %zu != %zu", size, sizeof(nd_uniforms));
    if (!out || !out->color) return nd_fail(c, ND_E_INVALID, "path trace needs a colour output");
    if (spp == 0 || spp > 65536) return nd_fail(c, ND_E_INVALID, "spp %u", spp);
    const nd_uniforms& u = *(const nd_uniforms*)uniforms;
    if (u.photon_enabled > 0.5f || u.fog_count > 0.5f || u.approx_flags != 0.0f)
        return nd_fail(c, ND_E_UNSUPPORTED, "photons, fog and approximations are not implemented");
    if (hipSetDevice(c->device) != hipSuccess) return nd_fail(c, ND_E_DEVICE, "hipSetDevice");
    nd_march_opts o{ND_NORMAL_TETRA, 0, 0, ND_KERNEL_PERSISTENT};
    DevJob job;
    std::memset(&job, 0, sizeof(job));
    nd_status st = nd_pack_job(c, u, o, rect, job);
    if (st != ND_OK) return st;
    if ((st = nd_accel_prepare(c, &job, 1)) != ND_OK) return st;
    nd::McScene S;
    pack_mc_scene(u, S);
    const uint32_t npix = (uint32_t)job.rect_w * (uint32_t)job.rect_h;
    // Sample chunk: the path pool holds up to 2^k paths (pixels x samples), k = mc_pool_log2, by
    // default the largest k in 24..27 whose pool fits in a quarter of the device's memory: on a
    // 288 GB device a 1080p frame of 64 samples is one chunk, with one tail of bounce passes. The
    // shadow queue (paths x lights entries of 8 floats) keeps 32-bit float indexes, so paths x
    // lights stays within 2^29. Adaptive sampling takes chunks of at most half its minimum sample
    // count, so that pixels converged in one chunk start no paths in the next.
    const size_t lights = (size_t)std::max(S.light_count, 1);
    const size_t slots = (size_t)S.slots;  // path slots per pixel sample (lens samples with depth of field)
    const bool dev_out = out->device_ptrs != 0;
    auto layout_bytes = [&](size_t P) {
        size_t bytes = 0;
        auto sz = [&](size_t b) { bytes += (b + 255) & ~(size_t)255; };
        sz(sizeof(DevJob)); sz(sizeof(DevScene)); sz(sizeof(nd::McScene));
        for (int i = 0; i < 6; ++i) sz(4 * P);  // pos, dir
        sz(32 * P); sz(32 * P);                 // thr, rad
        for (int i = 0; i < 4; ++i) sz(4 * P);  // hero, seeds, hit_t
        sz(P); sz(P);                           // depth, flags
        sz(4 * P); sz(2 * P);                   // lambda, inner
        sz(4 * (size_t)npix); sz(4 * (size_t)npix); sz(4 * (size_t)npix);  // primary hit cache
        sz(4 * (size_t)npix);                                               // depth-of-field aperture
        sz(4 * P); sz(4 * P);                   // live queues
        const size_t SQ = P * lights;
        sz(4 * SQ); for (int i = 0; i < 8; ++i) sz(4 * SQ); sz(32 * SQ);
        sz(4 * P); for (int i = 0; i < 7; ++i) sz(4 * P); sz(P); sz(32 * P);
        sz(256);
        sz(sizeof(nd::McChunk) * (nd::MC_MAX_CHUNKS + 1));
        if (!dev_out) { sz(12 * (size_t)npix); sz(4 * (size_t)npix); sz(12 * (size_t)npix); }
        return bytes;
    };
    uint32_t pool_log2 = c->tuning.mc_pool_log2;
    if (!pool_log2) {
        pool_log2 = 24;
        while (pool_log2 < 27 && layout_bytes((size_t)1 << (pool_log2 + 1)) <= c->mem_total / 4) ++pool_log2;
    }
    const size_t pool = std::min((size_t)1 << pool_log2, ((size_t)1 << 29) / lights);
    uint32_t chunk = (uint32_t)std::max<size_t>(1, std::min<size_t>(spp, pool / std::max<size_t>((size_t)npix * slots, 1)));
    if (S.adapt_on) chunk = std::min(chunk, std::max(1u, (uint32_t)(S.adapt_min_samples * 0.5f)));
    const size_t P = (size_t)npix * chunk * slots;
    if (P >= ((size_t)1 << 32)) return nd_fail(c, ND_E_INVALID, "%u pixels x %zu lens slots exceed 2^32 paths", npix, slots);
    if (P * lights > ((size_t)1 << 29)) return nd_fail(c, ND_E_INVALID, "%u pixels x %zu lights exceed the shadow queue's 2^29 entries", npix, lights);
    const size_t bytes = layout_bytes(P);
    const size_t SQ = P * lights;
    if ((st = c->mc.ensure(c, bytes)) != ND_OK) return st;
    char* base = (char*)c->mc.p;
    auto carve = [&](size_t b) { char* r = base; base += (b + 255) & ~(size_t)255; return r; };
    DevJob* d_job = (DevJob*)carve(sizeof(DevJob));
    DevScene* d_scene = (DevScene*)carve(sizeof(DevScene));
    nd::McScene* d_S = (nd::McScene*)carve(sizeof(nd::McScene));
    nd::McArgs a;
    std::memset(&a, 0, sizeof(a));
    nd::McPaths& Pp = a.P;
    Pp.px = (float*)carve(4 * P); Pp.py = (float*)carve(4 * P); Pp.pz = (float*)carve(4 * P);
    Pp.dx = (float*)carve(4 * P); Pp.dy = (float*)carve(4 * P); Pp.dz = (float*)carve(4 * P);
    Pp.thr = (float*)carve(32 * P); Pp.rad = (float*)carve(32 * P);
    Pp.hero = (float*)carve(4 * P); Pp.seedx = (float*)carve(4 * P); Pp.seedy = (float*)carve(4 * P);
    Pp.hit_t = (float*)carve(4 * P);
    Pp.depth = (uint8_t*)carve(P); Pp.flags = (uint8_t*)carve(P);
    Pp.lambda = (float*)carve(4 * P); Pp.inner = (uint16_t*)carve(2 * P);
    Pp.ct = (float*)carve(4 * (size_t)npix); Pp.cf = (uint32_t*)carve(4 * (size_t)npix); Pp.cn = (uint32_t*)carve(4 * (size_t)npix);
    Pp.dof_ap = (float*)carve(4 * (size_t)npix);
    Pp.live_a = (uint32_t*)carve(4 * P); Pp.live_b = (uint32_t*)carve(4 * P);
    Pp.shadow_path = (uint32_t*)carve(4 * SQ);
    Pp.sox = (float*)carve(4 * SQ); Pp.soy = (float*)carve(4 * SQ); Pp.soz = (float*)carve(4 * SQ);
    Pp.sdx = (float*)carve(4 * SQ); Pp.sdy = (float*)carve(4 * SQ); Pp.sdz = (float*)carve(4 * SQ);
    Pp.sdist = (float*)carve(4 * SQ); Pp.sk = (float*)carve(4 * SQ);
    Pp.sw = (float*)carve(32 * SQ);
    Pp.ind_path = (uint32_t*)carve(4 * P);
    Pp.iox = (float*)carve(4 * P); Pp.ioy = (float*)carve(4 * P); Pp.ioz = (float*)carve(4 * P);
    Pp.idx = (float*)carve(4 * P); Pp.idy = (float*)carve(4 * P); Pp.idz = (float*)carve(4 * P);
    Pp.it = (float*)carve(4 * P);
    Pp.ihit = (uint8_t*)carve(P);
    Pp.iw = (float*)carve(32 * P);
    Pp.ctr = (uint32_t*)carve(256);
    nd::McChunk* d_chunks = (nd::McChunk*)carve(sizeof(nd::McChunk) * (nd::MC_MAX_CHUNKS + 1));
    float* d_rgb = dev_out ? out->color : (float*)carve(12 * (size_t)npix);
    float* d_depth = dev_out ? out->depth : (out->depth ? (float*)carve(4 * (size_t)npix) : nullptr);
    float* d_normal = dev_out ? out->normal : (out->normal ? (float*)carve(12 * (size_t)npix) : nullptr);
    hipStream_t s = c->stream;
    hipError_t e;
    if ((e = hipMemcpyAsync(d_job, &job, sizeof(job), hipMemcpyHostToDevice, s)) != hipSuccess) return nd_hip_fail(c, e, "upload job");
    if ((e = hipMemcpyAsync(d_scene, &job.scene, sizeof(DevScene), hipMemcpyHostToDevice, s)) != hipSuccess) return nd_hip_fail(c, e, "upload scene");
    if ((e = hipMemcpyAsync(d_S, &S, sizeof(S), hipMemcpyHostToDevice, s)) != hipSuccess) return nd_hip_fail(c, e, "upload materials");
    a.job = d_job; a.scene = d_scene; a.S = d_S; a.ch = d_chunks + nd::MC_MAX_CHUNKS;
    a.npix = npix;
    a.chunk = 64;  // bounce queues are short: small claims balance better
    a.out_rgb = d_rgb; a.out_depth = d_depth; a.out_normal = d_normal;
    // each pass at its kernel's residency; claims of 64 paths over the pool
    const uint64_t claims = ((uint64_t)P + 63) / 64;
    const bool ext = nd_scene_ext(job.scene);
    const int rx = ext ? RES_MC_EXTEND_EXT - RES_MC_EXTEND : 0;
    const uint32_t g_ext = nd_pass_grid(c, RES_MC_EXTEND + rx, claims), g_shade = nd_pass_grid(c, RES_MC_SHADE + rx, claims);
    const uint32_t g_shadow = nd_pass_grid(c, RES_MC_SHADOW + rx, claims), g_im = nd_pass_grid(c, RES_MC_IND_MARCH + rx, claims);
    const uint32_t g_is = nd_pass_grid(c, RES_MC_IND_SHADE + rx, claims);
    const int max_depth = S.max_depth;
    const bool indirect = S.indirect_on != 0;
    const bool inside = S.transmissive != 0;  // paths can enter solids: extend runs march_inside
    const bool transp = S.transp != 0;        // shadow rays cross transparent occluders
    const bool adapt = S.adapt_on != 0;       // adaptive sampling: per-pixel state in c->mc_adapt
    const bool sss = S.sss != 0;              // subsurface materials: the shade pass's SSS variant
    const bool dof = S.dof_on != 0;           // depth of field: lens samples per out-of-focus pixel
    if (adapt) {
        if ((st = c->mc_adapt.ensure(c, 32 * (size_t)npix)) != ND_OK) return st;
        a.adapt = (float*)c->mc_adapt.p;
        const uint32_t key[6] = {(uint32_t)job.rect_x, (uint32_t)job.rect_y, (uint32_t)job.rect_w, (uint32_t)job.rect_h,
                                 (uint32_t)job.target_w, (uint32_t)job.target_h};
        // a new sequence (mc_sample 0) or another rect starts from empty statistics
        if (u.mc_sample < 0.5f || std::memcmp(key, c->mc_adapt_key, sizeof(key)) != 0) {
            hipLaunchKernelGGL(nd::k_mc_zero, dim3(256), dim3(256), 0, s, a.adapt, (uint64_t)8 * npix);
            std::memcpy(c->mc_adapt_key, key, sizeof(key));
        }
    }
    // The passes of one chunk, in stream order. Captured once into a graph per shape and
    // replayed for every chunk (the chunk's sample range is read from a.ch).
    auto enqueue_chunk = [&](hipStream_t st) {
        hipLaunchKernelGGL(nd::k_mc_counters, dim3(1), dim3(64), 0, st, Pp.ctr, 0);
        // with adaptive sampling or depth of field the generate pass appends the paths it starts
        hipLaunchKernelGGL(nd::k_mc_set_live, dim3(1), dim3(64), 0, st, Pp.ctr, (const nd::McChunk*)a.ch, npix,
                           (adapt || dof) ? 1 : 0);
        hipLaunchKernelGGL(nd::k_mc_generate, dim3((uint32_t)(((size_t)npix * chunk + 255) / 256)), dim3(256), 0, st, a);
        int ping = 0;
        for (int depth = 0; depth < max_depth; ++depth) {
            hipLaunchKernelGGL(nd::k_mc_counters, dim3(1), dim3(64), 0, st, Pp.ctr, 1);
            if (ext) {
                if (inside) hipLaunchKernelGGL((nd::k_mc_extend<true, true>), dim3(g_ext), dim3(64), 0, st, a, ping, a.scene, a.S);
                else hipLaunchKernelGGL((nd::k_mc_extend<true, false>), dim3(g_ext), dim3(64), 0, st, a, ping, a.scene, a.S);
                if (sss) hipLaunchKernelGGL((nd::k_mc_shade<true, true>), dim3(g_shade), dim3(64), 0, st, a, ping, a.scene, a.S);
                else hipLaunchKernelGGL((nd::k_mc_shade<true, false>), dim3(g_shade), dim3(64), 0, st, a, ping, a.scene, a.S);
                if (transp) hipLaunchKernelGGL((nd::k_mc_shadow<true, true>), dim3(g_shadow), dim3(64), 0, st, a, a.scene, a.S);
                else hipLaunchKernelGGL((nd::k_mc_shadow<true, false>), dim3(g_shadow), dim3(64), 0, st, a, a.scene, a.S);
            } else {
                if (inside) hipLaunchKernelGGL((nd::k_mc_extend<false, true>), dim3(g_ext), dim3(64), 0, st, a, ping, a.scene, a.S);
                else hipLaunchKernelGGL((nd::k_mc_extend<false, false>), dim3(g_ext), dim3(64), 0, st, a, ping, a.scene, a.S);
                if (sss) hipLaunchKernelGGL((nd::k_mc_shade<false, true>), dim3(g_shade), dim3(64), 0, st, a, ping, a.scene, a.S);
                else hipLaunchKernelGGL((nd::k_mc_shade<false, false>), dim3(g_shade), dim3(64), 0, st, a, ping, a.scene, a.S);
                if (transp) hipLaunchKernelGGL((nd::k_mc_shadow<false, true>), dim3(g_shadow), dim3(64), 0, st, a, a.scene, a.S);
                else hipLaunchKernelGGL((nd::k_mc_shadow<false, false>), dim3(g_shadow), dim3(64), 0, st, a, a.scene, a.S);
            }
            if (depth == 0 && indirect) {
                if (ext) {
                    hipLaunchKernelGGL(nd::k_mc_indirect_march<true>, dim3(g_im), dim3(64), 0, st, a, a.scene, a.S);
                    hipLaunchKernelGGL(nd::k_mc_indirect_shade<true>, dim3(g_is), dim3(64), 0, st, a, a.scene, a.S);
                } else {
                    hipLaunchKernelGGL(nd::k_mc_indirect_march<false>, dim3(g_im), dim3(64), 0, st, a, a.scene, a.S);
                    hipLaunchKernelGGL(nd::k_mc_indirect_shade<false>, dim3(g_is), dim3(64), 0, st, a, a.scene, a.S);
                }
            }
            hipLaunchKernelGGL(nd::k_mc_counters, dim3(1), dim3(64), 0, st, Pp.ctr, 2);
            ping ^= 1;
        }
        if (adapt) hipLaunchKernelGGL(nd::k_mc_resolve_adapt, dim3((npix + 255) / 256), dim3(256), 0, st, a);
        else hipLaunchKernelGGL(nd::k_mc_resolve, dim3((npix + 255) / 256), dim3(256), 0, st, a);
    };
    const bool use_graph = !(c->tuning.flags & ND_TUNE_NO_GRAPH);
    if (use_graph) {
        // everything the captured launches bake in: kernel arguments and launch shapes
        struct McGraphKey {
            nd::McArgs args;
            uint32_t grid[5];
            int max_depth, indirect, ext, inside, transp, adapt, sss, dof, slots;
            uint64_t paths;
        } key;
        static_assert(sizeof(McGraphKey) <= sizeof(c->mc_graph_key), "graph key storage");
        std::memset(&key, 0, sizeof(key));
        key.args = a;
        key.grid[0] = g_ext; key.grid[1] = g_shade; key.grid[2] = g_shadow; key.grid[3] = g_im; key.grid[4] = g_is;
        key.ext = ext;
        key.inside = inside;
        key.transp = transp;
        key.adapt = adapt;
        key.sss = sss;
        key.dof = dof;
        key.slots = (int)slots;
        key.max_depth = max_depth; key.indirect = indirect; key.paths = (uint64_t)P;
        if (!c->mc_graph_exec || std::memcmp(&key, c->mc_graph_key, sizeof(key)) != 0) {
            if (c->mc_graph_exec) { (void)hipGraphExecDestroy(c->mc_graph_exec); c->mc_graph_exec = nullptr; }
            hipGraph_t g = nullptr;
            if ((e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal)) != hipSuccess) return nd_hip_fail(c, e, "begin graph capture");
            enqueue_chunk(s);
            hipError_t le = hipGetLastError();
            if ((e = hipStreamEndCapture(s, &g)) != hipSuccess) return nd_hip_fail(c, e, "end graph capture");
            if (le != hipSuccess) { (void)hipGraphDestroy(g); return nd_hip_fail(c, le, "capture path-trace passes"); }
            e = hipGraphInstantiate(&c->mc_graph_exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) { c->mc_graph_exec = nullptr; return nd_hip_fail(c, e, "instantiate path-trace graph"); }
            std::memcpy(c->mc_graph_key, &key, sizeof(key));
        }
    }
    // chunk records, uploaded MC_MAX_CHUNKS at a time. With the primary hit cache, sample 0
    // runs as a chunk of its own so that every later chunk reuses its primary segments.
    std::vector<nd::McChunk> hc;
    for (uint32_t base = 0; base < spp;) {
        const uint32_t n = (S.cache_on && base == 0) ? 1u : std::min(chunk, spp - base);
        hc.push_back(nd::McChunk{base, n, spp, u.mc_sample});
        base += n;
    }
    const uint32_t n_chunks = (uint32_t)hc.size();
    hipLaunchKernelGGL(nd::k_mc_zero, dim3(1024), dim3(256), 0, s, d_rgb, (uint64_t)3 * npix);
    hipLaunchKernelGGL(nd::k_mc_zero, dim3(256), dim3(256), 0, s, (float*)Pp.cf, (uint64_t)npix);  // empty hit cache
    hipLaunchKernelGGL(nd::k_mc_counters, dim3(1), dim3(64), 0, s, Pp.ctr, 3);
    (void)hipEventRecord(c->ev[6], s);
    if (dof) {
        // each pixel's centre ray, once per call: the effective aperture of its hit
        const uint32_t g_probe = nd_pass_grid(c, RES_MC_EXTEND + rx, ((uint64_t)npix + 63) / 64);
        if (ext) hipLaunchKernelGGL(nd::k_mc_dof_probe<true>, dim3(g_probe), dim3(64), 0, s, a, a.scene, a.S);
        else hipLaunchKernelGGL(nd::k_mc_dof_probe<false>, dim3(g_probe), dim3(64), 0, s, a, a.scene, a.S);
    }
    for (uint32_t k0 = 0; k0 < n_chunks; k0 += nd::MC_MAX_CHUNKS) {
        const uint32_t nk = std::min(nd::MC_MAX_CHUNKS, n_chunks - k0);
        if (k0 > 0 && (e = hipStreamSynchronize(s)) != hipSuccess) return nd_hip_fail(c, e, "path trace execution");
        if ((e = hipMemcpyAsync(d_chunks, &hc[k0], sizeof(nd::McChunk) * nk, hipMemcpyHostToDevice, s)) != hipSuccess)
            return nd_hip_fail(c, e, "upload chunk records");
        for (uint32_t k = 0; k < nk; ++k) {
            // the passes read chunk k0 + k through a.ch: a one-thread kernel points it there
            hipLaunchKernelGGL(nd::k_mc_select_chunk, dim3(1), dim3(64), 0, s, d_chunks, k);
            if (use_graph) {
                if ((e = hipGraphLaunch(c->mc_graph_exec, s)) != hipSuccess) return nd_hip_fail(c, e, "launch path-trace graph");
            } else {
                enqueue_chunk(s);
            }
            if ((e = hipGetLastError()) != hipSuccess) return nd_hip_fail(c, e, "launch path trace");
            if (c->cancel && *c->cancel) {
                (void)hipStreamSynchronize(s);
                return nd_fail(c, ND_E_CANCELLED, "cancelled");
            }
        }
    }
    unsigned long long segs_total = 0, de_total = 0, shadows_total = 0;
    {
        uint32_t h_ctr[16];
        if ((e = hipMemcpyAsync(h_ctr, Pp.ctr, sizeof(h_ctr), hipMemcpyDeviceToHost, s)) != hipSuccess) return nd_hip_fail(c, e, "read counters");
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return nd_hip_fail(c, e, "path trace execution");
        segs_total = h_ctr[nd::C_SEG];
        de_total = *(unsigned long long*)(h_ctr + nd::C_DE);
        shadows_total = h_ctr[nd::C_SHTOT];
    }
    (void)hipEventRecord(c->ev[7], s);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return nd_hip_fail(c, e, "path trace execution");
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev[6], c->ev[7]);
    if (!dev_out) {
        if ((e = hipMemcpy(out->color, d_rgb, 12 * (size_t)npix, hipMemcpyDeviceToHost)) != hipSuccess) return nd_hip_fail(c, e, "read colour");
        if (out->depth && (e = hipMemcpy(out->depth, d_depth, 4 * (size_t)npix, hipMemcpyDeviceToHost)) != hipSuccess) return nd_hip_fail(c, e, "read depth");
        if (out->normal && (e = hipMemcpy(out->normal, d_normal, 12 * (size_t)npix, hipMemcpyDeviceToHost)) != hipSuccess) return nd_hip_fail(c, e, "read normal");
    }
    if (stats) {
        stats->paths = (uint64_t)npix * spp;
        stats->segments = segs_total;
        stats->shadow_rays = shadows_total;
        stats->de_evals = de_total;
        stats->ms_total = ms;
        stats->max_depth_reached = (uint32_t)S.max_depth;
    }
    return ND_OK;
}
