// nd_oracle_mc.hpp - TEST INFRASTRUCTURE ONLY. Scalar restatement of the reference's Monte
// Carlo path trace (trace_path and what it calls) for opaque scenes: spectral hero
// wavelengths, the procedural sky, point / global lights with soft shadows, Blinn-Phong
// shading, Schlick reflection with roughness-perturbed normals, the single cosine-sampled
// indirect bounce at the primary hit, Russian roulette. Citations: template.wgsl lines.
//
// Random numbers: the reference hashes the raw FP32 bits of screen-space seeds (hash12w,
// template.wgsl:2420-2434), so every seed here is computed in float exactly as the shader
// computes it. The oracle (T = double) then draws the same random numbers as an FP32
// renderer, and per-sample differences come from geometry rounding only.
#pragma once

#include "nd_oracle.hpp"

namespace ndo {

// ----------------------------------------------------------------------------- hashing
// hash_u32 / hash12w (template.wgsl:2420-2434)
inline uint32_t hash_u32(float px, float py) {
    uint32_t x, y;
    std::memcpy(&x, &px, 4);
    std::memcpy(&y, &py, 4);
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
inline float hash12w(float px, float py) { return (float)hash_u32(px, py) * 2.3283064365386963e-10f; }
// interleaved gradient noise (hash12, template.wgsl:2413-2415), in FP32
inline float ffract(float x) { return x - std::floor(x); }
inline float hash12(float px, float py) {
    const float d = px * 0.06711056f + py * 0.00583715f;
    return ffract(52.9829189f * ffract(d));
}
struct Seed { float x, y; };
inline Seed seed_add(Seed s, float ax, float ay) { return {s.x + ax, s.y + ay}; }

// ----------------------------------------------------------------------------- spectra
// 8 hero-stratified wavelengths and the RGB <-> spectrum bases (template.wgsl:713-883).
template <class T> struct Spec { T v[8]; };
template <class T> Spec<T> splat(T a) { Spec<T> s; for (auto& x : s.v) x = a; return s; }
template <class T> Spec<T> operator+(const Spec<T>& a, const Spec<T>& b) { Spec<T> r; for (int k = 0; k < 8; ++k) r.v[k] = a.v[k] + b.v[k]; return r; }
template <class T> Spec<T> operator*(const Spec<T>& a, const Spec<T>& b) { Spec<T> r; for (int k = 0; k < 8; ++k) r.v[k] = a.v[k] * b.v[k]; return r; }
template <class T> Spec<T> operator*(const Spec<T>& a, T s) { Spec<T> r; for (int k = 0; k < 8; ++k) r.v[k] = a.v[k] * s; return r; }
template <class T> T smax(const Spec<T>& a) { T m = a.v[0]; for (int k = 1; k < 8; ++k) m = std::max(m, a.v[k]); return m; }

template <class T> struct Spectral {
    T lambda[8];
    Spec<T> br, bg, bb;  // basis of R, G, B
    Spec<T> wr, wg, wb;  // projection weights back to R, G, B

    explicit Spectral(T hero) {
        const T lo = T(380.0), span = T(730.0 - 380.0);
        const double strata[8] = {0.0, 0.125, 0.25, 0.375, 0.5, 0.625, 0.75, 0.875};
        T base = tfract(hero);
        for (int k = 0; k < 8; ++k) lambda[k] = lo + span * tfract(base + T(strata[k]));
        const double FR[3] = {0.99575300, -0.00656071, 0.01080772};
        const double FG[3] = {-0.00755455, 1.00830220, -0.00074766};
        const double FB[3] = {0.02732810, -0.00674214, 0.97941404};
        T sr = T(0), sg = T(0), sb = T(0);
        for (int k = 0; k < 8; ++k) {
            T l = lambda[k];
            // spectral_prebasis (template.wgsl:789-797)
            T blue = T(1) / (T(1) + exp((l - T(498.09)) / T(13.50)));
            T red = T(1) / (T(1) + exp((T(600.31) - l) / T(8.00)));
            T t = (l - T(530.0)) / T(25.0);
            T green = T(2.49) * exp(T(-0.5) * t * t);
            T sum = red + green + blue;
            T a = red / sum, b = green / sum, c = blue / sum;
            br.v[k] = a * T(FR[0]) + b * T(FG[0]) + c * T(FB[0]);
            bg.v[k] = a * T(FR[1]) + b * T(FG[1]) + c * T(FB[1]);
            bb.v[k] = a * T(FR[2]) + b * T(FG[2]) + c * T(FB[2]);
            // spectral_cie (template.wgsl:805-814): piecewise Gaussian fits of x, y, z
            auto lobe = [&](double mu, double slo, double shi) {
                T s = l < T(mu) ? T(slo) : T(shi);
                T u = (l - T(mu)) / s;
                return exp(T(-0.5) * u * u);
            };
            T x = T(1.056) * lobe(599.8, 37.9, 31.0) + T(0.362) * lobe(442.0, 16.0, 26.7) - T(0.065) * lobe(501.1, 20.4, 26.2);
            T y = T(0.821) * lobe(568.8, 46.9, 40.5) + T(0.286) * lobe(530.9, 16.3, 31.1);
            T z = T(1.217) * lobe(437.0, 11.8, 36.0) + T(0.681) * lobe(459.0, 26.0, 13.8);
            wr.v[k] = T(3.2404542) * x + T(-1.5371385) * y + T(-0.4985314) * z;
            wg.v[k] = T(-0.9692660) * x + T(1.8760108) * y + T(0.0415560) * z;
            wb.v[k] = T(0.0556434) * x + T(-0.2040259) * y + T(1.0572252) * z;
            sr = sr + wr.v[k]; sg = sg + wg.v[k]; sb = sb + wb.v[k];
        }
        for (int k = 0; k < 8; ++k) { wr.v[k] = wr.v[k] / sr; wg.v[k] = wg.v[k] / sg; wb.v[k] = wb.v[k] / sb; }
    }
    // spec_from_rgb (template.wgsl:873-877): clamped at 0 per wavelength
    Spec<T> from_rgb(T r, T g, T b) const {
        Spec<T> s;
        for (int k = 0; k < 8; ++k) s.v[k] = tmax(r * br.v[k] + g * bg.v[k] + b * bb.v[k], T(0));
        return s;
    }
    void to_rgb(const Spec<T>& s, T* rgb) const {
        rgb[0] = rgb[1] = rgb[2] = T(0);
        for (int k = 0; k < 8; ++k) {
            rgb[0] = rgb[0] + s.v[k] * wr.v[k];
            rgb[1] = rgb[1] + s.v[k] * wg.v[k];
            rgb[2] = rgb[2] + s.v[k] * wb.v[k];
        }
    }
};

// gamut_fit (template.wgsl:898-908)
template <class T> void gamut_fit(T* c) {
    T y = c[0] * T(0.2126) + c[1] * T(0.7152) + c[2] * T(0.0722);
    if (y <= T(0)) { c[0] = c[1] = c[2] = T(0); return; }
    T m = tmin(c[0], tmin(c[1], c[2]));
    if (m >= T(0)) return;
    T f = y / (y - m);
    for (int k = 0; k < 3; ++k) c[k] = y + (c[k] - y) * f;
}

// ----------------------------------------------------------------------------- sky
// sky_color for the procedural sky (template.wgsl:2050-2094)
template <class T> V3<T> sky_color(const nd_uniforms& u, V3<T> dir) {
    if (((uint32_t)u.debug_parts & ND_PART_SKY) != 0u) return V3<T>();
    T c = T(std::cos((double)u.sky.yaw)), s = T(std::sin((double)u.sky.yaw));
    V3<T> d(c * dir.x - s * dir.z, dir.y, s * dir.x + c * dir.z);
    V3<T> base;
    V3<T> hz = V3<T>::from(u.sky.horizon), ze = V3<T>::from(u.sky.zenith), gr = V3<T>::from(u.sky.ground);
    if (d.y >= T(0)) {
        T t = pow(tclamp(d.y, T(0), T(1)), T(1.0 / std::max((double)u.sky.falloff, 0.01)));
        base = hz + (ze - hz) * t;
    } else {
        base = hz + (gr - hz) * tclamp(-d.y, T(0), T(1));
    }
    if (u.sky.sun_cos > -1.0f) {
        V3<T> sd = normalize(V3<T>::from(u.sky.sun_dir));
        T dd = dot(d, sd);
        T e0 = T(u.sky.sun_cos), e1 = T(u.sky.sun_cos) + (T(1) - T(u.sky.sun_cos)) * T(0.15);
        T x = tclamp((dd - e0) / (e1 - e0), T(0), T(1));
        T ss = x * x * (T(3) - T(2) * x);
        base = base + V3<T>::from(u.sky.sun_color) * (T(u.sky.sun_intensity) * ss);
    }
    T k = T(u.sky.intensity);
    return V3<T>(tmax(base.x * k, T(0)), tmax(base.y * k, T(0)), tmax(base.z * k, T(0)));
}

// ----------------------------------------------------------------------------- materials
struct Mat {
    double color[3];
    double glossiness, transparency, reflectiveness, ior, roughness, abbe, inner_max_steps;
    double film_thickness, film_ior, film_strength, film_angle_scale, film_perturb, film_perturb_scale;
    double subsurface, subsurface_radius;
};
inline Mat stop_mat(const nd_color_stop& c) {
    return Mat{{c.color[0], c.color[1], c.color[2]}, c.glossiness, c.transparency, c.reflectiveness, c.ior, c.roughness,
               c.abbe, c.inner_max_steps, c.film_thickness, c.film_ior, c.film_strength, c.film_angle_scale,
               c.film_perturb, c.film_perturb_scale, c.subsurface, c.subsurface_radius};
}
// default_material (template.wgsl:1965-1978), DEFAULT_SSS_RADIUS = 0.1
inline Mat default_mat() {
    return Mat{{0.8, 0.8, 0.85}, 0.3, 0.0, 0.0, 1.5, 0.0, 0.0, 64.0, 400.0, 1.8, 0.0, 1.0, 0.0, 0.2, 0.0, 0.1};
}
// apply_part_mask (template.wgsl:394-402) for the parts this path reads
inline Mat apply_part_mask(Mat m, uint32_t parts) {
    if (parts & ND_PART_COLOR_STRIPS) { m.color[0] = 0.72; m.color[1] = 0.72; m.color[2] = 0.73; }  // SIMPLE_ALBEDO
    if (parts & ND_PART_REFLECTIONS) m.reflectiveness = 0.0;
    if (parts & ND_PART_REFRACTION) m.transparency = 0.0;
    if (parts & ND_PART_SUBSURFACE) m.subsurface = 0.0;
    if (parts & ND_PART_IRIDESCENCE) m.film_strength = 0.0;
    if (parts & ND_PART_DISPERSION) m.abbe = 0.0;
    return m;
}
// mix_material(a, b, h) = mix(b, a, h) per field (template.wgsl:1980-2001)
inline Mat mix_mat(const Mat& a, const Mat& b, double h) {
    Mat r;
    for (int k = 0; k < 3; ++k) r.color[k] = b.color[k] * (1 - h) + a.color[k] * h;
    r.glossiness = b.glossiness * (1 - h) + a.glossiness * h;
    r.transparency = b.transparency * (1 - h) + a.transparency * h;
    r.reflectiveness = b.reflectiveness * (1 - h) + a.reflectiveness * h;
    r.ior = b.ior * (1 - h) + a.ior * h;
    r.roughness = b.roughness * (1 - h) + a.roughness * h;
    r.abbe = b.abbe * (1 - h) + a.abbe * h;
    r.inner_max_steps = b.inner_max_steps * (1 - h) + a.inner_max_steps * h;
    r.film_thickness = b.film_thickness * (1 - h) + a.film_thickness * h;
    r.film_ior = b.film_ior * (1 - h) + a.film_ior * h;
    r.film_strength = b.film_strength * (1 - h) + a.film_strength * h;
    r.film_angle_scale = b.film_angle_scale * (1 - h) + a.film_angle_scale * h;
    r.film_perturb = b.film_perturb * (1 - h) + a.film_perturb * h;
    r.film_perturb_scale = b.film_perturb_scale * (1 - h) + a.film_perturb_scale * h;
    r.subsurface = b.subsurface * (1 - h) + a.subsurface * h;
    r.subsurface_radius = b.subsurface_radius * (1 - h) + a.subsurface_radius * h;
    return r;
}
// sample_gradient (template.wgsl:2003-2023)
inline Mat sample_gradient(const nd_fractal_instance& fi, double t) {
    int count = std::min(i32(fi.color_count), ND_MAX_COLOR_STOPS);
    if (count <= 0) return default_mat();
    if (count == 1 || t <= fi.colors[0].position) return stop_mat(fi.colors[0]);
    if (t >= fi.colors[count - 1].position) return stop_mat(fi.colors[count - 1]);
    for (int i = 0; i < count - 1; ++i) {
        const nd_color_stop &a = fi.colors[i], &b = fi.colors[i + 1];
        if (t >= a.position && t <= b.position) {
            double f = (t - a.position) / std::max((double)(b.position - a.position), 1e-4);
            return mix_mat(stop_mat(a), stop_mat(b), 1.0 - f);
        }
    }
    return stop_mat(fi.colors[count - 1]);
}
// orbit_trap_to_strip (template.wgsl:2025-2035)
inline double trap_to_strip(const nd_fractal_instance& fi, double v) {
    double x = v / std::max((double)fi.trap_span, 1e-4) + fi.trap_offset;
    int rep = i32(fi.trap_repeat + 0.5f);
    if (rep == 1) return x - std::floor(x);
    if (rep == 2) { double f = x * 0.5 - std::floor(x * 0.5); return 1.0 - std::fabs(f * 2.0 - 1.0); }
    return std::min(std::max(x, 0.0), 1.0);
}
// hit_material (template.wgsl:2128-2155): every visible instance's strip material, blended
// by the combine weight.
template <class T, class P> Mat hit_material(const nd_uniforms& u, const Scene<P>& s, V3<T> pos) {
    bool first = true;
    T d = T(s.max_dist * 2.0 + 1.0);
    Mat mat = default_mat();
    for (int i = 0; i < s.count; ++i) {
        const Instance<P>& in = s.inst[i];
        if (in.hidden) continue;
        T di, ti;
        instance_de(in, pos, &di, &ti);
        Mat mi = sample_gradient(u.instances[i], trap_to_strip(u.instances[i], (double)val(ti)));
        if (first) { d = di; mat = mi; first = false; continue; }
        T nd, h;
        combine_de(in.combine, d, di, T(in.blend_k), &nd, &h);
        mat = mix_mat(mat, mi, (double)val(h));
        d = nd;
    }
    return mat;
}

// ----------------------------------------------------------------------------- shading
template <class T> V3<T> reflect3(V3<T> i, V3<T> n) { return i - n * (T(2) * dot(n, i)); }

// cosine_hemisphere_from (template.wgsl:3341-3350)
template <class T> V3<T> cosine_hemisphere(V3<T> n, float u1, float u2) {
    T r = sqrt(T(u1));
    T phi = T(u2) * T(6.28318530718);
    V3<T> local(r * cos(phi), r * sin(phi), sqrt(tmax(T(1) - T(u1), T(0))));
    V3<T> up = abs(n.z) > T(0.99) ? V3<T>(T(1), T(0), T(0)) : V3<T>(T(0), T(0), T(1));
    V3<T> tg = normalize(cross(up, n));
    V3<T> bt = cross(n, tg);
    return normalize(tg * local.x + bt * local.y + n * local.z);
}
template <class T> V3<T> cosine_sample(V3<T> n, Seed s) {
    return cosine_hemisphere(n, hash12w(s.x, s.y), hash12w(s.x + 91.73f, s.y + 31.13f));
}

// perturb_normal (template.wgsl:3396-3401): the normal bent toward a cosine-distributed
// direction by alpha = roughness^2; unchanged at alpha <= MIN_ROUGHNESS (1e-4).
template <class T> V3<T> perturb_normal(V3<T> n, T alpha, Seed s) {
    if (alpha <= T(1e-4)) return n;
    V3<T> h = cosine_sample(n, s);
    return normalize(n + (h - n) * alpha);
}

// Russian roulette of path_shade (template.wgsl:4107-4115): survive with probability
// clamp(max throughput, RR_MIN_SURVIVAL = 0.05, 1), decided by hash12w(seed + (59.3, 23.9));
// a survivor's throughput is scaled by 1 / survive (*scale). Returns false when the path ends.
template <class T> bool roulette(T thr_max, Seed sd, T* scale) {
    T survive = tclamp(thr_max, T(0.05), T(1));
    Seed rs = seed_add(sd, 59.3f, 23.9f);
    if (T(hash12w(rs.x, rs.y)) > survive) { *scale = T(0); return false; }
    *scale = T(1) / survive;
    return true;
}

// light_terms (template.wgsl:3076-3087): n.l clamped at 0 and the Blinn-Phong specular
// max(n.h, 0)^spec_power * glossiness with h = normalize(l + v), or 0 with the specular part off.
template <class T> void light_terms(V3<T> ld, V3<T> n, V3<T> view, T spec_power, T glossiness, bool specular_on,
                                    T* ndotl, T* spec) {
    *ndotl = tmax(dot(n, ld), T(0));
    V3<T> hd = normalize(ld + view);
    *spec = specular_on ? pow(tmax(dot(n, hd), T(0)), spec_power) * glossiness : T(0);
}

struct LightSampleD { double dir[3], color[3], dist; };
// light_sample (template.wgsl:2178-2229) at a surface point, i.e. light_sample_ex with a
// zero-length segment: point, global, and ray lights. A ray light is a Gaussian beam from its
// position along axis = ray_direction / max(|ray_direction|, 1e-6): at depth t0 along the axis
// and distance r from it, colour * brightness * exp(-r^2 / w^2) / w^2 with w = waist +
// max(t0, 0) tan(clamp(spread, 0, 1.5)), and nothing behind the source (t0 <= 0).
template <class T> void light_sample(const nd_light& L, double max_dist, V3<T> pos, V3<T>* dir, V3<T>* color, T* dist) {
    int type = i32(L.light_type + 0.5f);
    V3<T> lc = V3<T>::from(L.color) * T(L.brightness);
    if (type == 1) {
        *dir = normalize(V3<T>::from(L.position_or_direction));
        *color = lc;
        *dist = T(max_dist * 2.0);
        return;
    }
    V3<T> to = V3<T>::from(L.position_or_direction) - pos;
    T d = length(to);
    *dir = to / tmax(d, T(1e-4));
    *dist = d;
    if (type != 2) {
        *color = lc / tmax(d * d, T(0.01));
        return;
    }
    const V3<T> rd = V3<T>::from(L.ray_direction);
    const V3<T> axis = rd / tmax(length(rd), T(1e-6));
    const V3<T> u0 = -to;
    const T t0 = dot(u0, axis);
    const V3<T> perp0 = u0 - axis * t0;
    const T w = T(L.waist) + tmax(t0, T(0)) * T(std::tan(std::min(std::max((double)L.spread, 0.0), 1.5)));
    const T profile = exp(-dot(perp0, perp0) / (w * w));
    *color = lc * (t0 > T(0) ? profile / (w * w) : T(0));
}

// visibility_march (template.wgsl:2282-2342) with shadow_factor's parameters (:2343-2375):
// an opaque occluder blocks the light; with transparency in the scene, a transparent one
// tints it by mix(1, colour, a) * a per crossing while the march steps through its inside.
template <class T, class P>
Spec<T> visibility_march(const nd_uniforms& u, const Scene<P>& s, const Spectral<T>& sp, V3<T> origin, V3<T> dir,
                         T light_dist, T k, int steps, int* evals) {
    const uint32_t parts = (uint32_t)u.debug_parts;
    const bool opaque_only = !(u.has_transparency > 0.5f && !(parts & ND_PART_REFRACTION));
    T t = T(0.004);  // SHADOW_BIAS as t0
    T res = T(1);
    Spec<T> tint = splat(T(1));
    bool in_solid = false;
    T step_floor = T(0.004);  // SHADOW_STEP_FLOOR
    for (int i = 0; i < steps; ++i) {
        if (t >= light_dist) break;
        V3<T> p = origin + dir * t;
        if (!in_solid && s.accel) {
            // accelerator skip (template.wgsl:2297-2303): only where it cannot lower res
            T skip = accel_skip(s.accel.get(), p);
            if (skip > T(0.001) && k * skip >= res * t) {
                t = t + skip;
                continue;
            }
        }
        T d = probed_de(s, p);
        if (evals) (*evals)++;
        if (in_solid) {
            if (d > T(0.001) * t + T(0.0005)) in_solid = false;  // SHADOW_EXIT_SLOPE, _BIAS
            t = t + tmax(abs(d), step_floor);
            step_floor = step_floor * T(1.15);  // SHADOW_FLOOR_GROWTH
            continue;
        }
        if (d < T(0.0005)) {  // SHADOW_HIT_EPS
            if (opaque_only) return splat(T(0));
            const Mat m = hit_material(u, s, p);  // unmasked, as the reference reads it here
            if (evals) (*evals)++;
            const T a = tclamp(T(m.transparency), T(0), T(1));
            if (a < T(0.01)) return splat(T(0));  // SHADOW_OPAQUE_ALPHA
            tint = tint * (sp.from_rgb(T(1) + (T(m.color[0]) - T(1)) * a, T(1) + (T(m.color[1]) - T(1)) * a,
                                       T(1) + (T(m.color[2]) - T(1)) * a) * a);
            if (smax(tint) < T(0.004)) return splat(T(0));  // SHADOW_TINT_CUT
            in_solid = true;
            t = t + step_floor;
            continue;
        }
        res = tmin(res, k * d / t);
        t = t + tmax(d, T(0.001));  // SHADOW_MIN_STEP
    }
    return tint * tclamp(res, T(0), T(1));
}

// ----------------------------------------------------------------------------- thin film
// hash33 (template.wgsl:3439-3450): three values in [0, 1) per integer lattice cell, from
// the cell's coordinates truncated to i32 and read as u32, in FP32 as the shader computes
inline void hash33(double px, double py, double pz, float out[3]) {
    uint32_t v[3] = {(uint32_t)(int32_t)px, (uint32_t)(int32_t)py, (uint32_t)(int32_t)pz};
    const uint32_t M = 1664525u, A = 1013904223u;  // LCG_MULTIPLIER, LCG_INCREMENT
    for (auto& x : v) x = x * M + A;
    v[0] += v[1] * v[2];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    for (auto& x : v) x ^= x >> 16;
    v[0] += v[1] * v[2];
    v[1] += v[2] * v[0];
    v[2] += v[0] * v[1];
    for (int k = 0; k < 3; ++k) out[k] = (float)v[k] * 2.3283064365386963e-10f;  // INV_U32_RANGE
}
// film_noise (template.wgsl:3452-3467): smoothstep-weighted trilinear value noise in [-1, 1]
template <class T> V3<T> film_noise(V3<T> p) {
    const T ix = floor(p.x), iy = floor(p.y), iz = floor(p.z);
    const T fx = p.x - ix, fy = p.y - iy, fz = p.z - iz;
    const T wx = fx * fx * (T(3) - T(2) * fx), wy = fy * fy * (T(3) - T(2) * fy), wz = fz * fz * (T(3) - T(2) * fz);
    float h[2][2][2][3];
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx)
                hash33((double)val(ix) + dx, (double)val(iy) + dy, (double)val(iz) + dz, h[dz][dy][dx]);
    T r[3];
    for (int k = 0; k < 3; ++k) {
        auto mx = [&](int dy, int dz) { return T(h[dz][dy][0][k]) * (T(1) - wx) + T(h[dz][dy][1][k]) * wx; };
        const T y0 = mx(0, 0) * (T(1) - wy) + mx(1, 0) * wy;
        const T y1 = mx(0, 1) * (T(1) - wy) + mx(1, 1) * wy;
        r[k] = (y0 * (T(1) - wz) + y1 * wz) * T(2) - T(1);
    }
    return V3<T>(r[0], r[1], r[2]);
}
// film_normal (template.wgsl:3469-3476)
template <class T> V3<T> film_normal(V3<T> n, V3<T> pos, const Mat& m) {
    const T amount = tclamp(T(m.film_perturb), T(0), T(1));
    if (amount <= T(0.001)) return n;  // MATERIAL_EPSILON
    const V3<T> bend = film_noise(pos / tmax(T(m.film_perturb_scale), T(1e-4)));
    return normalize(n + bend * amount);
}
// thin_film_tint (template.wgsl:3478-3519): Airy reflectance of a film of index film_ior and
// thickness film_thickness (nm) on a base of index ior, s and p averaged, per wavelength,
// normalised by its peak and blended with 1 by film_strength
template <class T> Spec<T> thin_film_tint(const Spectral<T>& sp, T cos_view, const Mat& m) {
    const T strength = tclamp(T(m.film_strength), T(0), T(1));
    const T theta = acos(tclamp(cos_view, T(0), T(1))) * tmax(T(m.film_angle_scale), T(0));
    const T c1 = cos(tmin(theta, T(1.5607)));  // FILM_MAX_ANGLE
    const T s1 = sqrt(tmax(T(1) - c1 * c1, T(0)));
    const T n2 = tmax(T(m.film_ior), T(1)), n3 = tmax(T(m.ior), T(1));
    const T s2 = s1 / n2, c2 = sqrt(tmax(T(1) - s2 * s2, T(0)));
    const T s3 = s1 / n3, c3 = sqrt(tmax(T(1) - s3 * s3, T(0)));
    const T r12s = (c1 - n2 * c2) / (c1 + n2 * c2);
    const T r12p = (n2 * c1 - c2) / (n2 * c1 + c2);
    const T r23s = (n2 * c2 - n3 * c3) / (n2 * c2 + n3 * c3);
    const T r23p = (n3 * c2 - n2 * c3) / (n3 * c2 + n2 * c3);
    auto peak_of = [](T a, T b) { T x = (abs(a) + abs(b)) / (T(1) + abs(a * b)); return x * x; };
    const T peak = T(0.5) * (peak_of(r12s, r23s) + peak_of(r12p, r23p));
    if (peak < T(1e-5)) return splat(T(1));  // FILM_MIN_PEAK
    auto airy = [](T a, T b, T cd) { T x = T(2) * a * b * cd; return (a * a + b * b + x) / (T(1) + a * a * b * b + x); };
    const T path = T(4) * T(3.14159265358979) * n2 * tmax(T(m.film_thickness), T(0)) * c2;
    Spec<T> r;
    for (int k = 0; k < 8; ++k) {
        const T cd = cos(path / sp.lambda[k]);
        const T R = T(0.5) * (airy(r12s, r23s, cd) + airy(r12p, r23p, cd));
        r.v[k] = T(1) + (R / peak - T(1)) * strength;
    }
    return r;
}

// ----------------------------------------------------------------------------- subsurface
// Subsurface / subsurface_setup (template.wgsl:3088-3114): the amount, the radius (at least
// SSS_MIN_RADIUS 0.0005), each wavelength's reach (its albedo relative to the peak, squared,
// at least SSS_REACH_FLOOR 0.05) and a blurred normal, the tetrahedral gradient at
// pos + n * radius/2 with taps radius/4 (SSS_BLUR_PER_RADIUS, SSS_BLUR_TAP_FRACTION)
template <class T> struct Subsurface {
    T amount, radius;
    Spec<T> reach;
    V3<T> blur_n;
};
template <class T, class P>
Subsurface<T> subsurface_setup(const Scene<P>& s, V3<T> pos, V3<T> n, const Mat& m, const Spec<T>& albedo, int* evals) {
    Subsurface<T> r;
    r.amount = tclamp(T(m.subsurface), T(0), T(1));
    r.radius = tmax(T(m.subsurface_radius), T(0.0005));
    r.reach = splat(T(0));
    r.blur_n = n;
    if (r.amount <= T(0.001)) {  // MATERIAL_EPSILON
        r.amount = T(0);
        return r;
    }
    const T peak = tmax(smax(albedo), T(1e-4));  // EPSILON
    for (int k = 0; k < 8; ++k) {
        const T rel = albedo.v[k] / peak;
        r.reach.v[k] = tmax(rel * rel, T(0.05));
    }
    const T blur = r.radius * T(0.5);
    const T e = blur * T(0.5);
    const V3<T> c = pos + n * blur;
    const double K[4][3] = {{1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {1, 1, 1}};
    V3<T> g;
    for (int i = 0; i < 4; ++i) {
        V3<T> k(T(K[i][0]), T(K[i][1]), T(K[i][2]));
        g = g + k * probed_de(s, c + k * e);
    }
    if (evals) *evals += 4;
    const T len = length(g);
    if (len > T(1e-12)) r.blur_n = g / len;  // EPSILON_TINY
    return r;
}
// diffuse_weight with subsurface_wrap (template.wgsl:3116-3134): per wavelength, n.l is
// blended toward the blurred normal's by the reach and wrapped by SSS_WRAP (0.75) * reach;
// light from behind (max(-n.l, 0)) is added, and the profile is mixed in by the amount
template <class T> Spec<T> diffuse_weight(const Subsurface<T>& s, T ndl, V3<T> light_dir) {
    if (s.amount <= T(0)) return splat(tmax(ndl, T(0)));
    const T ndl_blur = dot(s.blur_n, light_dir);
    Spec<T> r;
    for (int k = 0; k < 8; ++k) {
        const T re = s.reach.v[k];
        const T w = T(0.75) * re;
        const T wrap = tmax(ndl + (ndl_blur - ndl) * re + w, T(0)) / (T(1) + w);
        r.v[k] = (T(1) - s.amount) * tmax(ndl, T(0)) + s.amount * (wrap + tmax(-ndl, T(0)));
    }
    return r;
}
// subsurface_transmit (template.wgsl:3136-3138): exp(-depth / (radius * reach)) per wavelength
template <class T> Spec<T> subsurface_transmit(const Subsurface<T>& s, T depth) {
    Spec<T> r;
    for (int k = 0; k < 8; ++k) r.v[k] = exp(-depth / (s.radius * s.reach.v[k]));
    return r;
}
// refine_inside_exit (template.wgsl:3047-3059): bisection of [inside_t, outside_t] on the
// inside-exit test DE > 0.0005 t + 0.00005, returning the outside end
template <class T, class P>
T refine_inside_exit(const Scene<P>& s, V3<T> origin, V3<T> dir, T lo, T hi, int iterations, int* evals) {
    for (int i = 0; i < iterations; ++i) {
        const T mid = T(0.5) * (lo + hi);
        const T d = probed_de(s, origin + dir * mid);
        if (evals) (*evals)++;
        if (d > T(0.0005) * mid + T(0.00005)) hi = mid;  // INSIDE_EXIT_SLOPE, INSIDE_EXIT_BIAS
        else lo = mid;
    }
    return hi;
}
// subsurface_light_path (template.wgsl:3140-3218): the transmission toward a light through
// subsurface material. From outside (n.l > 0) without occlusion it is 1. Otherwise the ray
// marches toward the light; inside a subsurface solid it takes steps of max(|DE|, floor),
// the floor starting at max(0.05 radius, 0.0005) and growing by 1 + 6/64, leaves when the
// inside-exit test holds after the warm-up (refined by 7 bisections), and returns 0 once
// the depth inside exceeds 6 radii. Outside it steps by max(DE, 0.001) with accelerator
// skips. A hit on a surface without subsurface hands the rest of the way to the soft
// shadow march (shadow_factor from 2 biases before the hit); without occlusion the march
// ends at the first exit. The result is exp(-depth / (radius * reach)) over the depth
// travelled inside, times that shadow.
template <class T, class P>
Spec<T> subsurface_light_path(const nd_uniforms& u, const Scene<P>& s, const Spectral<T>& sp, V3<T> pos, V3<T> n,
                              bool start_inside, const Subsurface<T>& ss, const nd_light& L, V3<T> ld, T ldist,
                              bool occlusion, int* evals) {
    if (!start_inside && !occlusion) return splat(T(1));
    const uint32_t parts = (uint32_t)u.debug_parts;
    const V3<T> origin = start_inside ? pos : pos + n * T(0.004);  // SHADOW_BIAS
    const T cap = ss.radius * T(6);                                // SSS_PROBE_DEPTH_RADII
    const T step_base = tmax(ss.radius * T(0.05), T(0.0005));      // SSS_PROBE_STEP_PER_RADIUS, INSIDE_MIN_STEP
    const T growth = T(1) + T(6) / T(64);                          // INSIDE_STEP_GROWTH / SSS_MAX_STEPS
    const int steps = std::min(i32(u.max_steps), 128);             // SSS_PATH_MAX_STEPS
    bool inside = start_inside;
    T t = T(0), prev = T(0), entry = T(0), depth = T(0), step_floor = step_base;
    int inner = 0;
    for (int i = 0; i < steps; ++i) {
        if (t >= ldist) break;
        const V3<T> p = origin + ld * t;
        if (!inside && s.accel) {
            const T skip = accel_skip(s.accel.get(), p);
            if (skip > T(0.001)) {  // SHADOW_MIN_STEP
                t = t + skip;
                continue;
            }
        }
        const T d = probed_de(s, p);
        if (evals) (*evals)++;
        if (inside) {
            if (inner > 1 && d > T(0.0005) * t + T(0.00005)) {  // INSIDE_WARMUP_STEPS, inside_exited
                const T exit_t = refine_inside_exit(s, origin, ld, prev, t, 7, evals);  // SSS_EXIT_REFINE_STEPS
                depth = depth + exit_t - entry;
                inside = false;
                if (!occlusion) break;
                t = exit_t + T(0.004);
                continue;
            }
            prev = t;
            t = t + tmax(abs(d), step_floor);
            step_floor = step_floor * growth;
            inner += 1;
            if (depth + t - entry > cap) return splat(T(0));
            continue;
        }
        if (d < T(0.0005)) {  // SHADOW_HIT_EPS
            const Mat hm = apply_part_mask(hit_material(u, s, p), parts);
            if (evals) (*evals)++;
            if (hm.subsurface <= 0.001) {
                // shadow_factor(p - 2 bias dir, dir, light, dir, ldist - t + 2 bias)
                Spec<T> rest = splat(T(1));
                if (L.cast_shadows > 0.5f) {
                    const double k = L.shadow_softness + (512.0 - L.shadow_softness) * L.hard_shadows;
                    rest = visibility_march(u, s, sp, p - ld * T(0.008) + ld * T(0.004), ld, ldist - t + T(0.008), T(k),
                                            std::min(i32(u.max_steps), 40), evals);
                }
                return subsurface_transmit(ss, depth) * rest;
            }
            inside = true;
            entry = t;
            prev = t;
            inner = 1;
            step_floor = step_base;
            t = t + step_floor;
            continue;
        }
        t = t + tmax(d, T(0.001));
    }
    if (inside) depth = depth + tmin(t, ldist) - entry;
    return subsurface_transmit(ss, depth);
}

// calc_ao (template.wgsl:1825-1841): 5 taps along the normal at radius (0.2 + 0.2 i), radius
// = clamp(0.04 t, 0.003, 0.25); occlusion = sum of max(tap - DE, 0) weighted 0.6^i; the factor
// clamp(1 - 2 occlusion / radius, 0, 1).
template <class T, class P> T calc_ao(const Scene<P>& s, V3<T> pos, V3<T> n, T t) {
    T radius = tclamp(t * T(0.04), T(0.003), T(0.25));
    T occ = T(0), w = T(1);
    for (int i = 0; i < 5; ++i) {
        T step = radius * (T(0.2) + T(0.2) * T((double)i));
        T dd = scene_de(s, pos + n * step);
        occ = occ + tmax(step - dd, T(0)) * w;
        w = w * T(0.6);
    }
    return tclamp(T(1) - T(2) * occ / radius, T(0), T(1));
}

// shade_local (template.wgsl:3222-3331), with the Monte Carlo light pick: when several
// shadowed lights reach the point and mc_light_pick is set, one of them is chosen in
// proportion to its unshadowed weight and scaled by total / weight. A subsurface material
// evaluates every light through subsurface_light_path with the wrapped diffuse weight.
template <class T, class P>
Spec<T> shade_local(const nd_uniforms& u, const Scene<P>& s, const Spectral<T>& sp, V3<T> pos, V3<T> dir, V3<T> n,
                    const Mat& m, const Spec<T>& ambient, bool use_shadows, Seed seed, int* evals,
                    Spec<T>* albedo_out = nullptr) {
    V3<T> view = -dir;
    T spec_power = T(4.0) + (T(128.0) - T(4.0)) * tclamp(T(m.glossiness), T(0), T(1));
    Spec<T> albedo = sp.from_rgb(T(m.color[0]), T(m.color[1]), T(m.color[2]));
    Spec<T> film = splat(T(1));
    if (m.film_strength > 0.001) {  // MATERIAL_EPSILON
        film = thin_film_tint(sp, dot(view, film_normal(n, pos, m)), m);
        albedo = albedo * film;
    }
    if (albedo_out) *albedo_out = albedo;  // g_shade_albedo
    const Subsurface<T> sss = subsurface_setup(s, pos, n, m, albedo, evals);
    Spec<T> diffuse = splat(T(0)), specular = splat(T(0));
    const uint32_t parts = (uint32_t)u.debug_parts;
    int lc = (parts & ND_PART_DIRECT_LIGHTS) ? 0 : std::min(i32(u.light_count), ND_MAX_LIGHTS);
    bool shadows_on = use_shadows && !(parts & ND_PART_SHADOWS);
    const bool pick_one = shadows_on && u.mc_enabled > 0.5f && u.mc_light_pick > 0.5f && lc > 1;
    T pick_w[ND_MAX_LIGHTS] = {};
    T pick_total = T(0);
    auto terms = [&](const nd_light& L, V3<T>* ld, V3<T>* col, T* dist, T* ndotl, T* spec) {
        light_sample(L, u.max_dist, pos, ld, col, dist);
        light_terms(*ld, n, view, spec_power, T(m.glossiness), !(parts & ND_PART_SPECULAR), ndotl, spec);
    };
    auto shadow_of = [&](const nd_light& L, V3<T> ld, T dist) {
        double k = L.shadow_softness + (512.0 - L.shadow_softness) * L.hard_shadows;
        int steps = std::min(i32(u.max_steps), 40);
        return visibility_march(u, s, sp, pos + n * T(0.004), ld, dist, T(k), steps, evals);
    };
    for (int i = 0; i < lc; ++i) {
        const nd_light& L = u.lights[i];
        V3<T> ld, col; T dist, ndotl, spec;
        terms(L, &ld, &col, &dist, &ndotl, &spec);
        if (tmax(col.x, tmax(col.y, col.z)) < T(1e-5)) continue;
        Spec<T> ls = sp.from_rgb(col.x, col.y, col.z);
        if (sss.amount > T(0)) {
            // subsurface: every light is evaluated here, through the material (no light pick)
            const T ndl = dot(n, ld);
            const Spec<T> w = diffuse_weight(sss, ndl, ld);
            const T sspec = ndl > T(0) ? spec : T(0);
            if (smax(w) <= T(0) && sspec <= T(0)) continue;
            const bool occlusion = shadows_on && L.cast_shadows > 0.5f;
            const Spec<T> reaching =
                ls * subsurface_light_path(u, s, sp, pos, n, ndl <= T(0), sss, L, ld, dist, occlusion, evals);
            diffuse = diffuse + reaching * w;
            specular = specular + reaching * sspec;
            continue;
        }
        if (ndotl <= T(0)) continue;
        if (pick_one && L.cast_shadows > 0.5f) {
            pick_w[i] = (col.x + col.y + col.z) * (ndotl + spec);
            pick_total = pick_total + pick_w[i];
            continue;
        }
        Spec<T> shadow = splat(T(1));
        if (shadows_on && L.cast_shadows > 0.5f) {
            shadow = shadow_of(L, ld, dist);
            if (smax(shadow) <= T(0)) continue;
        }
        Spec<T> reaching = ls * shadow;
        diffuse = diffuse + reaching * ndotl;
        specular = specular + reaching * spec;
    }
    if (pick_total > T(0)) {
        const T at = T(hash12w(seed.x + 47.23f, seed.y + 113.91f)) * pick_total;  // LIGHT_PICK_SEED_OFFSET
        int pick = 0;
        T running = T(0);
        for (int i = 0; i < lc; ++i) {
            if (pick_w[i] <= T(0)) continue;
            pick = i;
            running = running + pick_w[i];
            if (running > at) break;
        }
        const nd_light& L = u.lights[pick];
        V3<T> ld, col; T dist, ndotl, spec;
        terms(L, &ld, &col, &dist, &ndotl, &spec);
        const T scale = pick_total / pick_w[pick];
        Spec<T> reaching = sp.from_rgb(col.x, col.y, col.z) * shadow_of(L, ld, dist) * scale;
        diffuse = diffuse + reaching * ndotl;
        specular = specular + reaching * spec;
    }
    return albedo * (ambient + diffuse) + specular * film;
}

// fresnel_schlick (template.wgsl:3424-3430)
template <class T> T fresnel_schlick(T cos_theta, T ior) {
    T r0root = (T(1) - ior) / (T(1) + ior);
    T r0 = r0root * r0root;
    T c = tclamp(T(1) - abs(cos_theta), T(0), T(1));
    T c2 = c * c;
    return r0 + (T(1) - r0) * c2 * c2 * c;
}

// undo_tonemap (template.wgsl:4139-4142): c / max(1 - c, 1 / 255) of c clamped to [0, 1]
template <class T> V3<T> undo_tonemap(V3<T> c) {
    const T k = T(1.0 / 255.0);
    c = V3<T>(tclamp(c.x, T(0), T(1)), tclamp(c.y, T(0), T(1)), tclamp(c.z, T(0), T(1)));
    return V3<T>(c.x / tmax(T(1) - c.x, k), c.y / tmax(T(1) - c.y, k), c.z / tmax(T(1) - c.z, k));
}

// fs_simple's colour attachment (template.wgsl:4492-4538) and fs_aov's albedo (:4853-4896)
// for a primary ray that the march ended at (hit, t). With PART_GEOMETRY_ONLY set,
// simple_geometry (:4492-4512): a key light normalize(-d + 0.45 up - 0.35 right) with
// ambient 0.18 on albedo (0.72, 0.72, 0.73), mixed toward the background (0.10, 0.11, 0.13)
// by 0.55 t / max_dist, through undo_tonemap; otherwise render_lite with shade_lite
// (:4424-4444): (1 - F) (colour (0.15 AO + sum L n.l) + sum L (n.h)^p glossiness), no shadows.
// The normal is the tetrahedral estimate_normal at e = eps(t). aov_albedo: the colour mixed
// toward white by max(reflectiveness, transparency); a miss shows the sky, clamped for the albedo.
template <class T, class P>
void gbuffer_color(const nd_uniforms& u, const Scene<P>& s, const MarchCfg& mc, V3<T> org, V3<T> dir, bool hit, T t,
                   T* rgb, T* alb) {
    const uint32_t parts = (uint32_t)u.debug_parts;
    const bool simple = (parts & ND_PART_GEOMETRY_ONLY) != 0u;
    const V3<T> bg(T(0.10), T(0.11), T(0.13));
    V3<T> color, albedo;
    if (!hit) {
        const V3<T> sky = sky_color(u, dir);
        color = simple ? undo_tonemap(bg) : sky;
        albedo = V3<T>(tclamp(sky.x, T(0), T(1)), tclamp(sky.y, T(0), T(1)), tclamp(sky.z, T(0), T(1)));
    } else {
        const V3<T> hp = org + dir * t;
        const V3<T> n = normal_tetra(s, hp, T(mc.eps_c) * t + T(mc.eps_f));
        const Mat m = apply_part_mask(hit_material(u, s, hp), parts);
        if (simple) {
            const V3<T> key = normalize(-dir + V3<T>::from(u.camera_up) * T(0.45) - V3<T>::from(u.camera_right) * T(0.35));
            const T shade = T(0.18) + (T(1) - T(0.18)) * tmax(dot(n, key), T(0));
            const T cue = tclamp(t / tmax(T(u.max_dist), T(1e-4)), T(0), T(1)) * T(0.55);
            const V3<T> a = V3<T>(T(0.72), T(0.72), T(0.73)) * shade;
            color = undo_tonemap(a + (bg - a) * cue);
        } else {
            const T ao = (parts & ND_PART_AMBIENT_OCCLUSION) ? T(1) : calc_ao(s, hp, n, t);
            const T fres = fresnel_schlick(dot(dir, n), tmax(T(m.ior), T(1)));
            const T power = T(4.0) + (T(128.0) - T(4.0)) * tclamp(T(m.glossiness), T(0), T(1));
            V3<T> dif(T(0.15) * ao, T(0.15) * ao, T(0.15) * ao), spc;
            const int lc = (parts & ND_PART_DIRECT_LIGHTS) ? 0 : std::min(i32(u.light_count), ND_MAX_LIGHTS);
            for (int i = 0; i < lc; ++i) {
                V3<T> ld, lcol; T dist;
                light_sample(u.lights[i], u.max_dist, hp, &ld, &lcol, &dist);
                const T ndotl = tmax(dot(n, ld), T(0));
                if (ndotl <= T(0)) continue;
                dif = dif + lcol * ndotl;
                if (!(parts & ND_PART_SPECULAR)) spc = spc + lcol * (pow(tmax(dot(n, normalize(ld - dir)), T(0)), power) * T(m.glossiness));
            }
            const V3<T> mc3(T(m.color[0]), T(m.color[1]), T(m.color[2]));
            color = (cmul(mc3, dif) + spc) * (T(1) - fres);
        }
        const T sp = tclamp(tmax(T(m.reflectiveness), T(m.transparency)), T(0), T(1));
        albedo = V3<T>(tclamp(T(m.color[0]) + (T(1) - T(m.color[0])) * sp, T(0), T(1)),
                       tclamp(T(m.color[1]) + (T(1) - T(m.color[1])) * sp, T(0), T(1)),
                       tclamp(T(m.color[2]) + (T(1) - T(m.color[2])) * sp, T(0), T(1)));
    }
    rgb[0] = color.x; rgb[1] = color.y; rgb[2] = color.z;
    alb[0] = albedo.x; alb[1] = albedo.y; alb[2] = albedo.z;
}

// cauchy_ior (template.wgsl:3403-3409): n(lambda) = n_d + B (1 / lambda^2 - 1 / lambda_d^2),
// B = (n_d - 1) / (V (1 / lambda_F^2 - 1 / lambda_C^2)), lambda in micrometres
template <class T> T cauchy_ior(T ior_d, T abbe, T lambda) {
    T b = (ior_d - T(1)) / (tmax(abbe, T(1)) * T(1.9099));
    return ior_d + b * (T(1) / (lambda * lambda) - T(2.89626));
}

// refract (WGSL builtin) and refract_safe (template.wgsl:3387-3394): the normal faces the
// incoming ray; total internal reflection returns the mirror direction
template <class T> V3<T> refract3(V3<T> i, V3<T> n, T eta) {
    T c = dot(n, i);
    T k = T(1) - eta * eta * (T(1) - c * c);
    if (k < T(0)) return V3<T>();
    return i * eta - n * (eta * c + sqrt(k));
}
template <class T> V3<T> refract_safe(V3<T> dir, V3<T> raw_n, T eta, bool* tir) {
    V3<T> n = dot(dir, raw_n) > T(0) ? -raw_n : raw_n;
    V3<T> r = refract3(dir, n, eta);
    *tir = !(dot(r, r) > T(1e-4));  // REFRACT_TIR_EPSILON
    return *tir ? reflect3(dir, n) : r;
}

// dispersion_angles2 / dispersion_weights (template.wgsl:3521-3548): each wavelength's
// refracted direction against the hero's, as 2 (1 - cos), turned into Gaussian weights in
// angle that sum to 8
template <class T>
Spec<T> dispersion_weights(const Spectral<T>& sp, V3<T> in_dir, V3<T> n, const Mat& m, bool inside, V3<T> hero_dir,
                           T sigma) {
    const T ior_d = tmax(T(m.ior), T(1));
    Spec<T> g;
    T total = T(0);
    const T kk = T(-0.5) / tmax(sigma * sigma, T(1e-12));
    for (int k = 0; k < 8; ++k) {
        T ior_k = tmax(cauchy_ior(ior_d, T(m.abbe), sp.lambda[k] * T(0.001)), T(1));
        T eta_k = inside ? ior_k : T(1) / ior_k;
        bool tir;
        V3<T> d_k = refract_safe(in_dir, n, eta_k, &tir);
        T a = T(2) * (T(1) - tclamp(dot(hero_dir, d_k), T(-1), T(1)));
        g.v[k] = exp(a * kk);
        total = total + g.v[k];
    }
    return g * (T(8) / tmax(total, T(1e-6)));  // EPSILON_FINE
}

// march_inside (template.wgsl:3015-3046): from a point just inside a solid, steps of
// max(|DE|, floor) with a floor growing by 1 + 6 / steps, until the DE shows the ray has
// left (DE > 0.0005 t + 0.00005, from the third sample on) or the step budget ends
template <class T, class P>
MarchResult<T> march_inside(const Scene<P>& s, V3<T> org, V3<T> dir, int max_inner_steps, T max_t, int* evals) {
    MarchResult<T> r{};
    T t = T(0);
    const int steps = std::max(max_inner_steps, 1);
    const T growth = T(1) + T(6) / T((double)steps);
    T step_floor = T(0.004);  // max(INSIDE_STEP_BASE, INSIDE_MIN_STEP)
    r.hit = false;
    for (int i = 0; i < steps; ++i) {
        T d = scene_de(s, org + dir * t);
        if (evals) (*evals)++;
        if (i > 1 && d > T(0.0005) * t + T(0.00005)) {
            r.hit = true;
            r.dist = t;
            return r;
        }
        t = t + tmax(abs(d), step_floor);
        step_floor = step_floor * growth;
        if (t > max_t) break;
    }
    r.dist = t;
    return r;
}

// ----------------------------------------------------------------------------- adaptive sampling
// adapt_update (template.wgsl:4749-4760), in FP32 as the shader keeps the statistics: the
// luminance of the tonemapped sample c / (c + 1) is accumulated with its square and the
// count; from max(min_samples, 2) samples on, the pixel is converged (stats[3] = 1) when the
// standard error sqrt(var / n) is at most threshold * (mean + 0.05)
inline void adapt_update(float* stats, const double* rgb, float min_samples, float threshold) {
    const float w[3] = {0.2126f, 0.7152f, 0.0722f};  // LUMA_WEIGHTS
    float l = 0.f;
    for (int k = 0; k < 3; ++k) { const float c = (float)rgb[k]; l += c / (c + 1.f) * w[k]; }
    stats[0] += l; stats[1] += l * l; stats[2] += 1.f;
    if (stats[2] >= std::max(min_samples, 2.f)) {  // ADAPT_MIN_VARIANCE_SAMPLES
        const float mean = stats[0] / stats[2];
        const float var = std::max(stats[1] / stats[2] - mean * mean, 0.f);
        const float err = std::sqrt(var / stats[2]);
        stats[3] = err <= threshold * (mean + 0.05f) ? 1.f : 0.f;  // ADAPT_DARK_FLOOR
    }
}

// ----------------------------------------------------------------------------- hit cache
// oct_encode / oct_decode (template.wgsl:3696-3718): a unit normal through the octahedral
// map packed as two 16-bit snorm values (pack2x16snorm), and back. The primary hit cache
// keeps normals in this form, so with the cache on the path uses the decoded normal.
template <class T> V3<T> oct_quantize(V3<T> n) {
    const T l1 = abs(n.x) + abs(n.y) + abs(n.z);
    T px = n.x / l1, py = n.y / l1;
    if (n.z < T(0)) {
        const T qx = (T(1) - abs(py)) * (px >= T(0) ? T(1) : T(-1));
        const T qy = (T(1) - abs(px)) * (py >= T(0) ? T(1) : T(-1));
        px = qx; py = qy;
    }
    auto snorm = [](T x) {  // pack2x16snorm then unpack2x16snorm of one component
        const double q = std::floor(0.5 + 32767.0 * std::min(1.0, std::max(-1.0, (double)val(x))));
        return T(std::max(q / 32767.0, -1.0));
    };
    px = snorm(px); py = snorm(py);
    V3<T> r(px, py, T(1) - abs(px) - abs(py));
    const T t = tmax(-r.z, T(0));
    r.x = r.x + (r.x >= T(0) ? -t : t);
    r.y = r.y + (r.y >= T(0) ? -t : t);
    return normalize(r);
}

// ----------------------------------------------------------------------------- trace_path
struct PathResult { double rgb[3]; double depth; double normal[3]; int evals; int vertices; };

// trace_path (template.wgsl:4126-4137) with path_begin / path_extend / path_shade
// (:3819-4124): opaque and transmissive materials (Schlick reflection / refraction through
// march_inside), Cauchy dispersion on the picked wavelength with the Gaussian angle
// weights or the one-wavelength collapse, thin-film tints, the light pick and the primary
// hit cache's packed normals and subsurface materials. No fog, particles, photons or approximations.
// lens: a depth-of-field lens sample, which runs without the primary hit cache
// (g_primary_slot = -1, template.wgsl:5172).
template <class T, class P>
PathResult trace_path(const nd_uniforms& u, const Scene<P>& s, const MarchCfg& mc, V3<T> org, V3<T> dir, Seed seed,
                      bool lens = false) {
    PathResult out{};
    const float ms = u.mc_sample;
    Seed st_seed = seed_add(seed, ms * 27.31f, ms * 45.17f);
    float hero_f = hash12w(seed.x + 19.31f, seed.y + 47.11f) + ms * 0.61803399f;
    Spectral<T> sp{T(hero_f)};
    Spec<T> radiance = splat(T(0)), throughput = splat(T(1));
    V3<T> pos = org, d = dir;
    bool reflected = false, in_solid = false;
    // the primary hit cache (hit_cache_load/store, template.wgsl:3720-3766): on when the host
    // sets primary_cache_key; the primary hit and the hit after a mirror reflection at depth 0
    // keep their normals octahedral-packed
    const bool cache = u.primary_cache_key > 0.5f && !lens;
    bool second = false;
    int inner_steps = 64;  // PATH_DEFAULT_INNER_STEPS
    T lambda = T(0);        // picked wavelength in micrometres, 0 before the first dispersive hit
    const int max_depth = std::min(std::max(i32(u.max_reflection_bounces) + 8, 1), 16);
    const uint32_t parts = (uint32_t)u.debug_parts;
    // radians(0.5 * photon_dispersion_soft)
    const T soft = T((double)(0.5f * u.photon_dispersion_soft) * 0.017453292519943295);
    for (int depth = 0; depth < max_depth; ++depth) {
        out.vertices = depth + 1;
        Seed sd = seed_add(st_seed, (float)depth * 17.77f + 3.1f, (float)depth * 41.13f + 7.7f);
        MarchResult<T> m = in_solid ? march_inside(s, pos, d, inner_steps, T(u.max_dist), &out.evals)
                                    : march_to(s, mc, pos, d, T(u.max_dist));
        if (!in_solid) out.evals += m.evals;
        if (!m.hit) {
            V3<T> sky = sky_color(u, d);
            radiance = radiance + throughput * sp.from_rgb(sky.x, sky.y, sky.z);
            break;
        }
        if (smax(throughput) < T(0.002)) break;
        V3<T> hp = pos + d * m.dist;
        V3<T> n = normal_tetra(s, hp, T(mc.eps_c) * m.dist + T(mc.eps_f));
        out.evals += 4;
        if (cache && !in_solid && (depth == 0 || (depth == 1 && second))) n = oct_quantize(n);  // path_normal
        Mat mat = apply_part_mask(hit_material(u, s, hp), parts);
        out.evals += 1;
        if (depth == 0) {
            out.depth = (double)val(m.dist);
            out.normal[0] = (double)val(n.x); out.normal[1] = (double)val(n.y); out.normal[2] = (double)val(n.z);
        }
        Spec<T> disp_w = splat(T(1));
        if (lambda == T(0) && mat.abbe > 0.5) {  // MIN_ABBE
            const float hx = hash12w(sd.x + 3.71f, sd.y + 8.13f);  // DISPERSION_PICK_SEED_OFFSET
            const int pick = std::min((int)(hx * 8.0f), 7);
            lambda = sp.lambda[pick] * T(0.001);
            if (soft <= T(0)) {
                disp_w = splat(T(0));
                disp_w.v[pick] = T(8);  // spec_collapse_mask
            }
        }
        // scatter_at with allow_diffuse = false (template.wgsl:3552-3616)
        V3<T> nf = dot(d, n) > T(0) ? -n : n;
        T ior = tmax(T(mat.ior), T(1));
        if (lambda > T(0) && mat.abbe > 0.5) ior = tmax(cauchy_ior(ior, T(mat.abbe), lambda), T(1));
        const T eta = in_solid ? ior : T(1) / ior;
        T rough = tclamp(T(mat.roughness), T(0), T(1));
        T alpha = rough * rough;
        V3<T> nr = perturb_normal(nf, alpha, seed_add(sd, 5.19f, 12.73f));
        T f = fresnel_schlick(dot(d, nr), ior);
        Spec<T> film = splat(T(1));
        if (!in_solid && mat.film_strength > 0.001) film = thin_film_tint(sp, dot(-d, film_normal(nr, hp, mat)), mat);
        T transparency = tclamp(T(mat.transparency), T(0), T(1));
        T refl = tclamp(T(mat.reflectiveness), T(0), T(1));
        T p_reflect = tclamp(refl + (T(1) - refl) * f, T(0), T(1));
        T rest = T(1) - p_reflect;
        T p_transmit = rest * transparency;
        T local_share = rest * (T(1) - transparency);
        T p_cont = p_reflect + p_transmit;
        if (local_share > T(0.001)) {  // MIN_LOCAL_SHARE
            Spec<T> ambient = splat(T(0.15));
            bool mc_indirect = u.mc_enabled > 0.5f && !(parts & ND_PART_MC_INDIRECT);
            if (depth == 0 && mc_indirect) {
                // indirect_light_traced (template.wgsl:3359-3377)
                Seed is = seed_add(sd, 13.13f, 71.71f);
                V3<T> idir = cosine_sample(n, is);
                V3<T> io = hp + n * T(0.004);
                MarchResult<T> b = march_to(s, mc, io, idir, T(u.max_dist));
                out.evals += b.evals;
                if (!b.hit) {
                    V3<T> sky = sky_color(u, idir);
                    ambient = sp.from_rgb(sky.x, sky.y, sky.z);
                } else {
                    V3<T> bp = io + idir * b.dist;
                    V3<T> bn = normal_tetra(s, bp, T(mc.eps_c) * b.dist + T(mc.eps_f));
                    Mat bm = apply_part_mask(hit_material(u, s, bp), parts);
                    bm.subsurface = 0.0;  // indirect hits are shaded without subsurface
                    out.evals += 5;
                    ambient = shade_local(u, s, sp, bp, idir, bn, bm, splat(T(0.15)), false, is, &out.evals);
                }
            } else if (!reflected) {
                if (!(parts & ND_PART_AMBIENT_OCCLUSION)) {
                    out.evals += 5;
                    ambient = ambient * calc_ao(s, hp, n, m.dist);
                }
            }
            Spec<T> lit = shade_local(u, s, sp, hp, d, n, mat, ambient, !reflected, sd, &out.evals);
            radiance = radiance + throughput * lit * local_share;
        }
        // the continuation: r = hash * p_cont picks reflection below p_reflect, else refraction
        const T r = T(hash12w(sd.x, sd.y)) * p_cont;
        int kind;  // 0 absorbed, 1 reflect, 2 transmit
        V3<T> nd;
        Spec<T> w;
        if (r < p_reflect) {
            nd = reflect3(d, nr);
            w = sp.from_rgb(T(mat.color[0]) + (T(1) - T(mat.color[0])) * f, T(mat.color[1]) + (T(1) - T(mat.color[1])) * f,
                            T(mat.color[2]) + (T(1) - T(mat.color[2])) * f) * film * p_cont;
            kind = 1;
        } else if (r < p_reflect + p_transmit) {
            bool tir;
            nd = refract_safe(d, nr, eta, &tir);
            if (soft > T(0) && lambda > T(0) && mat.abbe > 0.5)
                disp_w = dispersion_weights(sp, d, nr, mat, in_solid, nd, soft);
            w = sp.from_rgb(T(mat.color[0]), T(mat.color[1]), T(mat.color[2])) * p_cont;
            kind = tir ? 1 : 2;
        } else {
            break;  // SCATTER_ABSORBED
        }
        if (kind == 1 && !in_solid && (parts & ND_PART_REFLECTIONS)) break;
        throughput = throughput * (disp_w * w);
        reflected = reflected || kind == 1;
        second = depth == 0 && !in_solid && kind == 1 && rough * rough <= T(1e-4);  // MIN_ROUGHNESS
        if (kind == 2) {
            in_solid = !in_solid;
            if (in_solid) inner_steps = std::min(std::max(i32((float)mat.inner_max_steps), 4), 512);
        }
        T bias = tmax(T(0.003), T(2) * (T(mc.eps_c) * m.dist + T(mc.eps_f)));
        pos = hp + nd * bias;
        d = nd;
        if (depth >= 4) {  // PATH_RR_START
            T scale;
            if (!roulette(smax(throughput), sd, &scale)) break;
            throughput = throughput * scale;
        }
    }
    T rgb[3];
    sp.to_rgb(radiance, rgb);
    gamut_fit(rgb);
    for (int k = 0; k < 3; ++k) out.rgb[k] = (double)val(rgb[k]);
    return out;
}

}  // namespace ndo
