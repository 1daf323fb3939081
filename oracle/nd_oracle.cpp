// nd_oracle.cpp - TEST INFRASTRUCTURE ONLY: C entry points over the scalar restatement in
// nd_oracle.hpp, loaded by tests/ (ctypes), the smoke check and the bench's cpu_baseline
// leg. Never linked into, or called by, the product library.
//
// prec = 0 evaluates in FP64 (the oracle), prec = 1 in FP32 (the CPU restatement).
#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "nd_oracle.hpp"
#include "nd_oracle_mc.hpp"

using namespace ndo;

namespace {

CodeMap code_map(const int32_t* cm) {
    CodeMap m;
    if (cm)
        for (int i = 0; i < 16; ++i) m.map[i] = cm[i];
    return m;
}

struct GbufOpts {
    int normal_mode;       // 0 analytic, 1 tetra, 2 central (h = 0.5 eps)
    int shadow_max_steps;  // 0: min(max_steps, 40)
    int shadows;
};

template <class T>
void pixel_gbuffer(const nd_uniforms& u, const Scene<T>& s, const MarchCfg& mc, const GbufOpts& o, int px,
                   int py, float* depth, float* normal, float* trap, uint8_t* hit, float* shadow,
                   uint32_t* evals) {
    V3<T> dir = primary_dir<T>(u, px, py, (int)u.resolution[0], (int)u.resolution[1]);
    V3<T> org = V3<T>::from(u.camera_pos) + V3<T>::from(s.cam_offset);
    MarchResult<T> m = march_to(s, mc, org, dir, T(u.max_dist));
    int ev = m.evals;
    if (!m.hit) {
        *depth = 0; normal[0] = normal[1] = normal[2] = 0; *trap = 0; *hit = 0; *shadow = 1;
        *evals = (uint32_t)ev;
        return;
    }
    V3<T> hp = org + dir * m.dist;
    T eps = T(mc.eps_c) * m.dist + T(mc.eps_f);
    V3<T> n;
    if (o.normal_mode == 1) n = normal_tetra(s, hp, eps);
    else if (o.normal_mode == 2) n = normal_central(s, hp, T(0.5) * eps);
    else {
        // analytic: forward-mode derivative evaluated in FP64 at the (T-precision) hit point
        const Scene<double> sd = recast_scene<double>(s);
        V3<double> nd = normal_analytic(sd, V3<double>((double)hp.x, (double)hp.y, (double)hp.z));
        n = V3<T>(T(nd.x), T(nd.y), T(nd.z));
    }
    T tr = T(0);
    scene_de(s, hp, &tr);
    float sh = 1.0f;
    if (o.shadows && (int)u.light_count >= 1) {
        const nd_light& L = u.lights[0];
        if (L.cast_shadows > 0.5f) {
            V3<T> ld; T dist;
            light_dir_dist(L, u.max_dist, hp, &ld, &dist);
            if (dot(n, ld) <= T(0)) {
                sh = 0.0f;  // shade_local skips a light with n.l <= 0 (template.wgsl:3276-3278)
            } else {
                double k = L.shadow_softness + (512.0 - L.shadow_softness) * L.hard_shadows;
                int steps = o.shadow_max_steps > 0 ? o.shadow_max_steps : std::min(i32(u.max_steps), 40);
                int sev = 0;
                T vis = visibility_march_opaque(s, hp + n * T(0.004), ld, dist, T(k), steps, &sev);
                ev += sev;
                sh = (float)val(vis);
            }
        }
    }
    *depth = (float)val(m.dist);
    normal[0] = (float)val(n.x); normal[1] = (float)val(n.y); normal[2] = (float)val(n.z);
    *trap = (float)val(tr);
    *hit = 1;
    *shadow = sh;
    *evals = (uint32_t)ev;
}

template <class T>
int render_gbuffer_t(const nd_uniforms* u, const int32_t* cm, const GbufOpts& o, int rx, int ry, int rw, int rh,
                     float* depth, float* normal, float* trap, uint8_t* hit, float* shadow, uint32_t* evals,
                     int nthreads, const double* cam_offset) {
    Scene<T> s = decode_scene<T>(*u, code_map(cm));
    attach_accel(*u, s);
    if (cam_offset) for (int k = 0; k < 3; ++k) s.cam_offset[k] = cam_offset[k];
    MarchCfg mc = march_cfg(*u);
    std::atomic<int> next_row{0};
    auto worker = [&]() {
        for (;;) {
            int y = next_row.fetch_add(1);
            if (y >= rh) break;
            for (int x = 0; x < rw; ++x) {
                size_t i = (size_t)y * rw + x;
                float dd, nn[3], tt, ss; uint8_t hh; uint32_t ee;
                pixel_gbuffer<T>(*u, s, mc, o, rx + x, ry + y, &dd, nn, &tt, &hh, &ss, &ee);
                if (depth) depth[i] = dd;
                if (normal) { normal[3 * i] = nn[0]; normal[3 * i + 1] = nn[1]; normal[3 * i + 2] = nn[2]; }
                if (trap) trap[i] = tt;
                if (hit) hit[i] = hh;
                if (shadow) shadow[i] = ss;
                if (evals) evals[i] = ee;
            }
        }
    };
    nthreads = std::max(1, nthreads);
    std::vector<std::thread> pool;
    for (int i = 1; i < nthreads; ++i) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    return 0;
}

}  // namespace

// The direct-lighting colour of a primary hit at depth t along d from o, as trace_path
// computes it at depth 0 with Monte Carlo off and the reflection continuation cut
// (PART_REFLECTIONS): the material of the hit point (constant over the surface in the scenes
// this is used on), the tetrahedral normal at eps(t), Schlick's split local_share =
// (1 - p_reflect)(1 - transparency), ambient 0.15 * AO, shade_local with shadows, then the
// spectrum to RGB and gamut_fit (template.wgsl:3819-4137). T is double, or Dual<6> with the
// scene's parameter tangents and o's camera tangents, for d(colour)/d(theta).
template <class T>
static void direct_rgb(const nd_uniforms& u, const Scene<T>& s, const Scene<double>& s64, const MarchCfg& mc, V3<T> o,
                       V3<double> d, T t, int px, int py, T* rgb) {
    const uint32_t parts = (uint32_t)u.debug_parts;
    const float ms = u.mc_sample;
    const Seed seed{((float)px + 0.5f) + ms * 91.73f, ((float)py + 0.5f) + ms * 57.31f};
    const Seed st_seed = seed_add(seed, ms * 27.31f, ms * 45.17f);
    const float hero_f = hash12w(seed.x + 19.31f, seed.y + 47.11f) + ms * 0.61803399f;
    const Seed sd = seed_add(st_seed, 3.1f, 7.7f);  // depth 0
    Spectral<T> sp{T(hero_f)};
    const V3<T> dir(T(d.x), T(d.y), T(d.z));
    const V3<T> hp = o + dir * t;
    const V3<T> n = normal_tetra(s, hp, T(mc.eps_c) * t + T(mc.eps_f));
    const V3<double> hp64(val(hp.x), val(hp.y), val(hp.z));
    const Mat mat = apply_part_mask(hit_material(u, s64, hp64), parts);
    const V3<T> nf = dot(dir, n) > T(0) ? -n : n;
    const T ior = tmax(T(mat.ior), T(1));
    const T f = fresnel_schlick(dot(dir, nf), ior);
    const T transparency = tclamp(T(mat.transparency), T(0), T(1));
    const T refl = tclamp(T(mat.reflectiveness), T(0), T(1));
    const T p_reflect = tclamp(refl + (T(1) - refl) * f, T(0), T(1));
    const T local_share = (T(1) - p_reflect) * (T(1) - transparency);
    Spec<T> radiance = splat(T(0));
    if (local_share > T(0.001)) {
        Spec<T> ambient = splat(T(0.15));
        if (!(parts & ND_PART_AMBIENT_OCCLUSION)) ambient = ambient * calc_ao(s, hp, n, t);
        int evals = 0;
        radiance = shade_local(u, s, sp, hp, dir, n, mat, ambient, true, sd, &evals) * local_share;
    }
    sp.to_rgb(radiance, rgb);
    gamut_fit(rgb);
}

extern "C" {

int ndo_abi_size(void) { return (int)sizeof(nd_uniforms); }

// image_sample (template.wgsl:1180-1193) of the image in `slot` at n points (u, v): 12
// doubles per point (value, near, far, 4 channels each).
int ndo_image_sample(int slot, const double* uv, int n, double* out) {
    if (slot < 0 || slot >= 16) return 1;
    const ndo::OImage& im = ndo::oracle_image(slot);
    for (int i = 0; i < n; ++i) {
        const ndo::OImageSample<double> s = ndo::image_sample(im, uv[2 * i], uv[2 * i + 1]);
        for (int c = 0; c < 4; ++c) {
            out[12 * i + c] = s.v[c];
            out[12 * i + 4 + c] = s.nr[c];
            out[12 * i + 8 + c] = s.fr[c];
        }
    }
    return 0;
}

// A height-map image for the heightmap formula in slot `slot` (0..15): n_words packed texel
// words (4 per texel, as nd_image_pack makes them), its size and per-channel slope bound.
// words null or w 0: the slot holds no image.
int ndo_bind_image(int slot, const uint32_t* words, int n_words, int w, int h, const double* slope) {
    if (slot < 0 || slot >= 16) return 1;
    ndo::OImage& im = ndo::oracle_image(slot);
    if (!words || w <= 0 || h <= 0) { im = ndo::OImage(); return 0; }
    if ((long)n_words != 4L * w * h) return 1;
    im.words.assign(words, words + n_words);
    im.w = (uint32_t)w;
    im.h = (uint32_t)h;
    for (int c = 0; c < 4; ++c) im.slope[c] = slope[c];
    return 0;
}

int ndo_scene_de(const nd_uniforms* u, const int32_t* cm, int prec, const double* p, int n, double* d_out,
                 double* trap_out) {
    if (prec == 0) {
        Scene<double> s = decode_scene<double>(*u, code_map(cm));
        for (int i = 0; i < n; ++i) {
            double tr = 0;
            d_out[i] = scene_de(s, V3<double>(p[3 * i], p[3 * i + 1], p[3 * i + 2]), trap_out ? &tr : nullptr);
            if (trap_out) trap_out[i] = tr;
        }
    } else {
        Scene<float> s = decode_scene<float>(*u, code_map(cm));
        for (int i = 0; i < n; ++i) {
            float tr = 0;
            d_out[i] = scene_de(s, V3<float>((float)p[3 * i], (float)p[3 * i + 1], (float)p[3 * i + 2]),
                                trap_out ? &tr : nullptr);
            if (trap_out) trap_out[i] = tr;
        }
    }
    return 0;
}

int ndo_march(const nd_uniforms* u, const int32_t* cm, int prec, const double* o, const double* d, int n,
              int32_t* hit, double* t, int32_t* steps, int32_t* evals) {
    MarchCfg mc = march_cfg(*u);
    if (prec == 0) {
        Scene<double> s = decode_scene<double>(*u, code_map(cm));
        attach_accel(*u, s);
        for (int i = 0; i < n; ++i) {
            auto m = march_to(s, mc, V3<double>(o[3 * i], o[3 * i + 1], o[3 * i + 2]),
                              V3<double>(d[3 * i], d[3 * i + 1], d[3 * i + 2]), (double)u->max_dist);
            hit[i] = m.hit; t[i] = m.dist; steps[i] = m.steps; evals[i] = m.evals;
        }
    } else {
        Scene<float> s = decode_scene<float>(*u, code_map(cm));
        attach_accel(*u, s);
        for (int i = 0; i < n; ++i) {
            auto m = march_to(s, mc, V3<float>((float)o[3 * i], (float)o[3 * i + 1], (float)o[3 * i + 2]),
                              V3<float>((float)d[3 * i], (float)d[3 * i + 1], (float)d[3 * i + 2]), u->max_dist);
            hit[i] = m.hit; t[i] = m.dist; steps[i] = m.steps; evals[i] = m.evals;
        }
    }
    return 0;
}

// mode: 0 analytic (FP64 dual numbers), 1 tetra e = eps, 2 central h = eps / 2; e is given.
// prec 1: the FP32 restatement of the finite-difference stencils (mode 1 tetrahedral, 2
// central); the analytic normal (mode 0) exists in FP64 only.
int ndo_normal(const nd_uniforms* u, const int32_t* cm, int prec, int mode, const double* pos, const double* e, int n,
               double* out) {
    if (prec != 0 && mode != 1 && mode != 2) return 1;
    Scene<double> s = decode_scene<double>(*u, code_map(cm));
    Scene<float> sf = decode_scene<float>(*u, code_map(cm));
    for (int i = 0; i < n; ++i) {
        V3<double> p(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
        V3<double> r;
        if (prec != 0) {
            V3<float> q((float)p.x, (float)p.y, (float)p.z);
            V3<float> rf = mode == 1 ? normal_tetra(sf, q, (float)e[i]) : normal_central(sf, q, (float)e[i]);
            r = V3<double>(rf.x, rf.y, rf.z);
        } else {
            r = mode == 1 ? normal_tetra(s, p, e[i]) : mode == 2 ? normal_central(s, p, e[i]) : normal_analytic(s, p);
        }
        out[3 * i] = r.x; out[3 * i + 1] = r.y; out[3 * i + 2] = r.z;
    }
    return 0;
}

// sky_color (template.wgsl:2050-2112) for n directions: RGB per direction.
int ndo_sky(const nd_uniforms* u, const double* dirs, int n, double* out) {
    for (int i = 0; i < n; ++i) {
        V3<double> c = sky_color(*u, V3<double>(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
        out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
    }
    return 0;
}

// calc_ao (template.wgsl:1825-1841) of the scene at n points with their normals and ray depths.
int ndo_calc_ao(const nd_uniforms* u, const int32_t* cm, const double* pos, const double* nrm, const double* t, int n,
                double* out) {
    Scene<double> s = decode_scene<double>(*u, code_map(cm));
    for (int i = 0; i < n; ++i)
        out[i] = calc_ao(s, V3<double>(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
                         normalize(V3<double>(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2])), t[i]);
    return 0;
}

int ndo_shadow(const nd_uniforms* u, const int32_t* cm, const double* origin, const double* dir, double light_dist,
               double k, int steps, double* vis, int32_t* evals) {
    Scene<double> s = decode_scene<double>(*u, code_map(cm));
    attach_accel(*u, s);
    int ev = 0;
    *vis = visibility_march_opaque(s, V3<double>(origin[0], origin[1], origin[2]), V3<double>(dir[0], dir[1], dir[2]),
                                   light_dist, k, steps, &ev);
    if (evals) *evals = ev;
    return 0;
}

int ndo_primary_dir(const nd_uniforms* u, int px, int py, double* out) {
    V3<double> d = primary_dir<double>(*u, px, py, (int)u->resolution[0], (int)u->resolution[1]);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
    return 0;
}

int ndo_render_gbuffer(const nd_uniforms* u, const int32_t* cm, int prec, int normal_mode, int shadow_max_steps,
                       int shadows, int rx, int ry, int rw, int rh, float* depth, float* normal, float* trap,
                       uint8_t* hit, float* shadow, uint32_t* evals, int nthreads, const double* cam_offset) {
    GbufOpts o{normal_mode, shadow_max_steps, shadows};
    if (prec == 0)
        return render_gbuffer_t<double>(u, cm, o, rx, ry, rw, rh, depth, normal, trap, hit, shadow, evals, nthreads,
                                        cam_offset);
    return render_gbuffer_t<float>(u, cm, o, rx, ry, rw, rh, depth, normal, trap, hit, shadow, evals, nthreads,
                                   cam_offset);
}

// fs_simple colour and fs_aov albedo of a rect (3 floats per pixel each; either may be null).
int ndo_gbuffer_color(const nd_uniforms* u, const int32_t* cm, int prec, int rx, int ry, int rw, int rh, float* color,
                      float* albedo, int nthreads) {
    auto run = [&](auto zero) {
        using T = decltype(zero);
        Scene<T> s = decode_scene<T>(*u, code_map(cm));
        attach_accel(*u, s);
        MarchCfg mc = march_cfg(*u);
        std::atomic<int> next_row{0};
        auto worker = [&]() {
            for (;;) {
                int y = next_row.fetch_add(1);
                if (y >= rh) break;
                for (int x = 0; x < rw; ++x) {
                    V3<T> dir = primary_dir<T>(*u, rx + x, ry + y, (int)u->resolution[0], (int)u->resolution[1]);
                    V3<T> org = V3<T>::from(u->camera_pos);
                    MarchResult<T> m = march_to(s, mc, org, dir, T(u->max_dist));
                    T rgb[3], alb[3];
                    gbuffer_color(*u, s, mc, org, dir, m.hit, m.dist, rgb, alb);
                    size_t i = (size_t)y * rw + x;
                    for (int k = 0; k < 3; ++k) {
                        if (color) color[3 * i + k] = (float)val(rgb[k]);
                        if (albedo) albedo[3 * i + k] = (float)val(alb[k]);
                    }
                }
            }
        };
        std::vector<std::thread> pool;
        for (int i = 1; i < std::max(1, nthreads); ++i) pool.emplace_back(worker);
        worker();
        for (auto& th : pool) th.join();
    };
    if (prec == 0) run(0.0); else run(0.0f);
    return 0;
}

// Mean ND-FLOP v1 per scene_de over the given points (uniform-only subexpressions
// hoisted); *tops_out receives the mean transcendental count.
double ndo_flops_per_de(const nd_uniforms* u, const int32_t* cm, const double* p, int n, double* tops_out) {
    Scene<double> s = decode_scene<double>(*u, code_map(cm));
    g_flops = FlopCounter{};
    for (int i = 0; i < n; ++i) {
        V3<Flop> q(Flop(p[3 * i]), Flop(p[3 * i + 1]), Flop(p[3 * i + 2]));
        (void)scene_de(s, q);
    }
    if (tops_out) *tops_out = n ? (double)g_flops.tops / n : 0.0;
    return n ? (double)g_flops.flops / n : 0.0;
}

// FP64 G-buffer render of a rect with the FLOP probe on: mean ND-FLOP v1 and T-ops per
// scene_de over the march (which = 0) or shadow (which = 1) evaluation points.
int ndo_flops_render(const nd_uniforms* u, const int32_t* cm, int shadow_max_steps, int rx, int ry, int rw, int rh,
                     int step, double* out) {
    Scene<double> s = decode_scene<double>(*u, code_map(cm));
    attach_accel(*u, s);
    MarchCfg mc = march_cfg(*u);
    uint64_t ev[2] = {0, 0}, fl[2] = {0, 0}, tp[2] = {0, 0};
    for (int y = 0; y < rh; y += step)
        for (int x = 0; x < rw; x += step) {
            // march evaluations
            g_probe = FlopProbe{};
            g_probe.on = true;
            V3<double> dir = primary_dir<double>(*u, rx + x, ry + y, (int)u->resolution[0], (int)u->resolution[1]);
            V3<double> org = V3<double>::from(u->camera_pos);
            MarchResult<double> m = march_to(s, mc, org, dir, (double)u->max_dist);
            ev[0] += g_probe.evals; fl[0] += g_probe.flops; tp[0] += g_probe.tops;
            g_probe = FlopProbe{};
            if (!m.hit || (int)u->light_count < 1 || u->lights[0].cast_shadows < 0.5f) continue;
            V3<double> hp = org + dir * m.dist;
            V3<double> n = normal_analytic(s, hp);
            V3<double> ld; double dist;
            light_dir_dist(u->lights[0], u->max_dist, hp, &ld, &dist);
            if (dot(n, ld) <= 0) continue;
            const nd_light& L = u->lights[0];
            double k = L.shadow_softness + (512.0 - L.shadow_softness) * L.hard_shadows;
            int steps = shadow_max_steps > 0 ? shadow_max_steps : std::min(i32(u->max_steps), 40);
            g_probe.on = true;
            (void)visibility_march_opaque(s, hp + n * 0.004, ld, dist, k, steps, (int*)nullptr);
            ev[1] += g_probe.evals; fl[1] += g_probe.flops; tp[1] += g_probe.tops;
            g_probe = FlopProbe{};
        }
    for (int w = 0; w < 2; ++w) {
        out[3 * w] = (double)ev[w];
        out[3 * w + 1] = ev[w] ? (double)fl[w] / ev[w] : 0.0;
        out[3 * w + 2] = ev[w] ? (double)tp[w] / ev[w] : 0.0;
    }
    return 0;
}

// FP64 reference gradients of the hit depth for a list of pixels.
// Parameters: 0 Mandelbulb power, 1 Mandelbox scale, 2 Mandelbox min radius (first visible
// instance of each formula), 3..5 camera translation. For each pixel:
//   t_root   the polished hit depth (0 = miss)
//   g_ift    dt/dp from the implicit-function theorem with FP64 dual numbers at t_root:
//            -(dDE/dp) / (grad DE . d - c), camera: -(grad DE) / (grad DE . d - c)
//   g_fd     central differences (t(p + h) - t(p - h)) / 2h of polished roots
//   ok       bit k set when both FD roots exist for parameter k
int ndo_adjoint_pixels(const nd_uniforms* u, const int32_t* cm, const int32_t* pix, int n, double h, double* t_root,
                       double* g_ift, double* g_fd, int32_t* ok) {
    using D6 = Dual<6>;
    const CodeMap map = code_map(cm);
    Scene<double> s = decode_scene<double>(*u, map);
    MarchCfg mc = march_cfg(*u);
    int bulb = -1, box = -1;
    for (int i = 0; i < s.count; ++i) {
        if (s.inst[i].hidden) continue;
        if (s.inst[i].formula == F_MANDELBULB && bulb < 0) bulb = i;
        if (s.inst[i].formula == F_MANDELBOX && box < 0) box = i;
    }
    Scene<D6> sd = decode_scene<D6>(*u, map);
    if (bulb >= 0) sd.inst[bulb].p.p[0] = D6::var(s.inst[bulb].p.p[0], 3);
    if (box >= 0) {
        sd.inst[box].p.p[0] = D6::var(s.inst[box].p.p[0], 4);
        sd.inst[box].p.p[1] = D6::var(s.inst[box].p.p[1], 5);
    }
    V3<double> o0 = V3<double>::from(u->camera_pos);
    auto perturbed = [&](int k, double delta, V3<double>* o) {
        Scene<double> q = s;
        *o = o0;
        if (k == 0 && bulb >= 0) q.inst[bulb].p.p[0] += delta;
        if (k == 1 && box >= 0) q.inst[box].p.p[0] += delta;
        if (k == 2 && box >= 0) q.inst[box].p.p[1] += delta;
        if (k == 3) o->x += delta;
        if (k == 4) o->y += delta;
        if (k == 5) o->z += delta;
        return q;
    };
    for (int i = 0; i < n; ++i) {
        int px = pix[2 * i], py = pix[2 * i + 1];
        V3<double> d = primary_dir<double>(*u, px, py, (int)u->resolution[0], (int)u->resolution[1]);
        double t;
        ok[i] = 0;
        for (int k = 0; k < 6; ++k) g_ift[6 * i + k] = g_fd[6 * i + k] = 0;
        if (!polished_root(s, mc, o0, d, &t)) { t_root[i] = 0; continue; }
        t_root[i] = t;
        V3<double> p = o0 + d * t;
        V3<D6> q(D6::var(p.x, 0), D6::var(p.y, 1), D6::var(p.z, 2));
        D6 de = scene_de(sd, q);
        double den = de.d[0] * d.x + de.d[1] * d.y + de.d[2] * d.z - mc.eps_c;
        if (den != 0) {
            g_ift[6 * i + 0] = -de.d[3] / den;
            g_ift[6 * i + 1] = -de.d[4] / den;
            g_ift[6 * i + 2] = -de.d[5] / den;
            g_ift[6 * i + 3] = -de.d[0] / den;
            g_ift[6 * i + 4] = -de.d[1] / den;
            g_ift[6 * i + 5] = -de.d[2] / den;
        }
        for (int k = 0; k < 6; ++k) {
            V3<double> op, om;
            Scene<double> sp = perturbed(k, h, &op), sm = perturbed(k, -h, &om);
            double tp, tm;
            if (polished_root(sp, mc, op, d, &tp) && polished_root(sm, mc, om, d, &tm)) {
                g_fd[6 * i + k] = (tp - tm) / (2 * h);
                ok[i] |= 1 << k;
            }
        }
    }
    return 0;
}

// Monte Carlo path trace of a rect, spp samples per pixel with mc_sample = u.mc_sample + s
// (the host's per-sample stamp, fractal_renderer.zig:3761-3775), averaged. Per-sample colors
// (spp x h x w x 3) are written too when sample_rgb is non-null.
// dof_aperture (template.wgsl:4726-4730): aperture * smoothstep(range / 2, range / 2 +
// max(4 aperture, 0.05), |hit_t - focus_t|)
static double dof_aperture(const nd_uniforms& u, double hit_t, double focus_t) {
    const double half = std::max((double)u.focus_range, 0.0) * 0.5;
    const double falloff = std::max((double)u.aperture * 4.0, 0.05);
    const double x = std::min(std::max((std::fabs(hit_t - focus_t) - half) / falloff, 0.0), 1.0);
    return (double)u.aperture * x * x * (3.0 - 2.0 * x);
}

int ndo_render_mc(const nd_uniforms* u0, const int32_t* cm, int prec, int rx, int ry, int rw, int rh, int spp,
                  float* rgb, float* depth, float* normal, float* sample_rgb, uint32_t* evals, int nthreads) {
    std::atomic<int> next_row{0};
    // depth of field: on with dof_enabled, aperture > DOF_MIN_APERTURE and PART_DEPTH_OF_FIELD;
    // lens samples min(mc_inner_count(32), the wavefront's inner slots)
    const bool dof = u0->dof_enabled > 0.5f && u0->aperture > 0.0001f && !((uint32_t)u0->debug_parts & ND_PART_DEPTH_OF_FIELD);
    const int inner = std::min(std::max((int)std::max(u0->mc_inner_samples, 1.0f), 1), 32);
    const int dof_n = std::min(u0->mc_enabled > 0.5f ? std::min(std::max((int)u0->mc_inner_samples, 1), 32) : 32, inner);
    Scene<double> scene64;
    Scene<float> scene32;
    if (prec == 0) { scene64 = decode_scene<double>(*u0, code_map(cm)); attach_accel(*u0, scene64); }
    else { scene32 = decode_scene<float>(*u0, code_map(cm)); attach_accel(*u0, scene32); }
    auto run = [&](auto tag) {
        using T = decltype(tag);
        nd_uniforms u = *u0;
        Scene<T> s;
        if constexpr (std::is_same<T, double>::value) s = scene64; else s = scene32;
        MarchCfg mc = march_cfg(u);
        for (;;) {
            int y = next_row.fetch_add(1);
            if (y >= rh) break;
            for (int x = 0; x < rw; ++x) {
                size_t i = (size_t)y * rw + x;
                double acc[3] = {0, 0, 0};
                uint32_t ev = 0;
                int px = rx + x, py = ry + y;
                V3<T> dir = primary_dir<T>(u, px, py, (int)u.resolution[0], (int)u.resolution[1]);
                V3<T> org = V3<T>::from(u.camera_pos);
                // depth of field (cs_wf_generate, template.wgsl:5150-5182): the centre ray's
                // hit sets the effective aperture; an out-of-focus pixel takes dof_n lens
                // samples per sample, whose gamut-fitted colours are averaged (fs_wf_resolve)
                int lens = 1;
                double ap = 0.0;
                V3<T> focus = org;
                if (dof) {
                    const V3<T> fw = V3<T>::from(u.camera_forward);
                    const T ft = T((double)u.focus_distance) / tmax(dot(dir, fw), T(1e-4));
                    focus = org + dir * ft;
                    MarchResult<T> pr = march_to(s, mc, org, dir, T(mc.max_dist));
                    const double probe_t = pr.hit ? (double)pr.dist : mc.max_dist;
                    ap = dof_aperture(*u0, probe_t, (double)ft);
                    if (ap > 0.0001) lens = dof_n;
                }
                // adaptive sampling (adapt_update, template.wgsl:4749-4760; the sample skip of
                // fs_main / cs_wf_generate): a converged pixel takes no more samples and
                // keeps the mean of those it took
                const bool adapt = u0->adapt_enabled > 0.5f && u0->mc_enabled > 0.5f && (float)px < u0->adapt_row_width;
                float st4[4] = {0.f, 0.f, 0.f, 0.f};
                int taken = 0;
                for (int sidx = 0; sidx < spp; ++sidx) {
                    u.mc_sample = u0->mc_sample + (float)sidx;
                    const float ms = u.mc_sample;
                    if (adapt && ms > 0.5f && st4[3] > 0.5f) continue;
                    Seed base{((float)px + 0.5f) + ms * 91.73f, ((float)py + 0.5f) + ms * 57.31f};
                    PathResult r;
                    if (lens == 1) {
                        r = trace_path<T>(u, s, mc, org, dir, base);
                    } else {
                        double sum[3] = {0, 0, 0};
                        for (int j = 0; j < lens; ++j) {
                            // seed + s (91.37, 51.91); sample_unit_disk (template.wgsl:4709-4713)
                            const Seed sj{base.x + (float)j * 91.37f, base.y + (float)j * 51.91f};
                            const double rr = std::sqrt((double)hash12(sj.x, sj.y));
                            const double th = (double)hash12w(sj.x + 17.17f, sj.y + 71.71f) * 6.28318530718;
                            const double dx = rr * std::cos(th) * ap, dy = rr * std::sin(th) * ap;
                            V3<T> o2(T((double)u.camera_pos[0] + u.camera_right[0] * dx + u.camera_up[0] * dy),
                                     T((double)u.camera_pos[1] + u.camera_right[1] * dx + u.camera_up[1] * dy),
                                     T((double)u.camera_pos[2] + u.camera_right[2] * dx + u.camera_up[2] * dy));
                            const V3<T> d2 = normalize(focus - o2);
                            const PathResult rj = trace_path<T>(u, s, mc, o2, d2, sj, true);
                            for (int k = 0; k < 3; ++k) sum[k] += rj.rgb[k];
                            if (j == 0) r = rj;
                            else r.evals += rj.evals;
                        }
                        for (int k = 0; k < 3; ++k) r.rgb[k] = sum[k] / lens;
                    }
                    ++taken;
                    if (adapt) adapt_update(st4, r.rgb, u0->adapt_min_samples, u0->adapt_threshold);
                    for (int k = 0; k < 3; ++k) acc[k] += r.rgb[k];
                    if (sample_rgb)
                        for (int k = 0; k < 3; ++k) sample_rgb[((size_t)sidx * rw * rh + i) * 3 + k] = (float)r.rgb[k];
                    if (sidx == 0) {
                        if (depth) depth[i] = (float)r.depth;
                        if (normal) for (int k = 0; k < 3; ++k) normal[3 * i + k] = (float)r.normal[k];
                    }
                    ev += (uint32_t)r.evals;
                }
                for (int k = 0; k < 3; ++k) rgb[3 * i + k] = (float)(acc[k] / std::max(taken, 1));
                if (evals) evals[i] = ev;
            }
        }
    };
    nthreads = std::max(1, nthreads);
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t) {
        if (prec == 0) pool.emplace_back([&] { run(0.0); });
        else pool.emplace_back([&] { run(0.0f); });
    }
    for (auto& th : pool) th.join();
    return 0;
}

// The direct-lighting colour (direct_rgb) of pixels and its derivative with respect to theta =
// (Mandelbulb power, Mandelbox scale, Mandelbox min radius, camera x, y, z): the hit depth's
// implicit-function tangent, then dual numbers through the normal, AO, lights and shadows.
// Without t_in the hit is the polished root of DE(o + d t) = eps(t) and g_fd gets central
// differences with step h (the perturbed scene's own polished root each time; ok bit k set
// where both roots exist); with t_in the dual is taken at the given depths (the GPU's). rgb
// gets the colour, g_dual and g_fd 18 values per pixel: [c * 6 + k] = d rgb_c / d theta_k.
int ndo_radiance_pixels(const nd_uniforms* u, const int32_t* cm, const int32_t* pix, const double* t_in, int n,
                        double h, double* t_root, double* rgb, double* g_dual, double* g_fd, int32_t* ok) {
    using D6 = Dual<6>;
    const CodeMap map = code_map(cm);
    Scene<double> s = decode_scene<double>(*u, map);
    MarchCfg mc = march_cfg(*u);
    int bulb = -1, box = -1;
    for (int i = 0; i < s.count; ++i) {
        if (s.inst[i].hidden) continue;
        if (s.inst[i].formula == F_MANDELBULB && bulb < 0) bulb = i;
        if (s.inst[i].formula == F_MANDELBOX && box < 0) box = i;
    }
    // dual slots: 0..2 camera x, y, z (and space, for the implicit-function tangent), 3 power,
    // 4 box scale, 5 min radius; outputs in theta order (power, scale, min radius, camera)
    Scene<D6> sd = decode_scene<D6>(*u, map);
    if (bulb >= 0) sd.inst[bulb].p.p[0] = D6::var(s.inst[bulb].p.p[0], 3);
    if (box >= 0) {
        sd.inst[box].p.p[0] = D6::var(s.inst[box].p.p[0], 4);
        sd.inst[box].p.p[1] = D6::var(s.inst[box].p.p[1], 5);
    }
    const int slot_of[6] = {3, 4, 5, 0, 1, 2};
    const V3<double> o0 = V3<double>::from(u->camera_pos);
    const int W = (int)u->resolution[0], H = (int)u->resolution[1];
    for (int i = 0; i < n; ++i) {
        const int px = pix[2 * i], py = pix[2 * i + 1];
        const V3<double> d = primary_dir<double>(*u, px, py, W, H);
        for (int k = 0; k < 18; ++k) g_dual[18 * i + k] = g_fd[18 * i + k] = 0;
        for (int c = 0; c < 3; ++c) rgb[3 * i + c] = 0;
        ok[i] = 0;
        double t;
        if (t_in) t = t_in[i];
        else if (!polished_root(s, mc, o0, d, &t)) t = 0;
        t_root[i] = t;
        if (!(t > 0)) continue;
        // dt/dtheta from one dual evaluation with spatial seeds (slots 0..2) and the parameters
        const V3<double> p = o0 + d * t;
        const D6 de = scene_de(sd, V3<D6>(D6::var(p.x, 0), D6::var(p.y, 1), D6::var(p.z, 2)));
        const double den = de.d[0] * d.x + de.d[1] * d.y + de.d[2] * d.z - mc.eps_c;
        if (den == 0) continue;
        D6 td(t);
        for (int k = 0; k < 6; ++k) td.d[k] = -de.d[k] / den;
        const V3<D6> od(D6::var(o0.x, 0), D6::var(o0.y, 1), D6::var(o0.z, 2));
        D6 c6[3];
        direct_rgb(*u, sd, s, mc, od, d, td, px, py, c6);
        for (int c = 0; c < 3; ++c) {
            rgb[3 * i + c] = c6[c].v;
            for (int k = 0; k < 6; ++k) g_dual[18 * i + 6 * c + k] = c6[c].d[slot_of[k]];
        }
        if (t_in) continue;
        for (int k = 0; k < 6; ++k) {
            double cp[3], cmn[3];
            bool both = true;
            for (int sgn = 0; sgn < 2 && both; ++sgn) {
                const double delta = sgn == 0 ? h : -h;
                Scene<double> q = s;
                V3<double> o = o0;
                if (k == 0 && bulb >= 0) q.inst[bulb].p.p[0] += delta;
                if (k == 1 && box >= 0) q.inst[box].p.p[0] += delta;
                if (k == 2 && box >= 0) q.inst[box].p.p[1] += delta;
                if (k == 3) o.x += delta;
                if (k == 4) o.y += delta;
                if (k == 5) o.z += delta;
                double tk;
                if (!polished_root(q, mc, o, d, &tk)) { both = false; break; }
                direct_rgb(*u, q, q, mc, o, d, tk, px, py, sgn == 0 ? cp : cmn);
            }
            if (!both) continue;
            ok[i] |= 1 << k;
            for (int c = 0; c < 3; ++c) g_fd[18 * i + 6 * c + k] = (cp[c] - cmn[c]) / (2 * h);
        }
    }
    return 0;
}

// FP64 implicit-function gradient at given hit points p = o + d t (t per pixel, e.g. the
// GPU's own hit depths), so the FP32 gradient can be compared at the point it was taken.
int ndo_adjoint_at(const nd_uniforms* u, const int32_t* cm, const int32_t* pix, const double* t_in, int n,
                   double* g_out) {
    using D6 = Dual<6>;
    const CodeMap map = code_map(cm);
    Scene<double> s = decode_scene<double>(*u, map);
    MarchCfg mc = march_cfg(*u);
    Scene<D6> sd = decode_scene<D6>(*u, map);
    bool bulb = false, box = false;
    for (int i = 0; i < sd.count; ++i) {
        if (sd.inst[i].hidden) continue;
        if (sd.inst[i].formula == F_MANDELBULB && !bulb) { sd.inst[i].p.p[0] = D6::var(s.inst[i].p.p[0], 3); bulb = true; }
        if (sd.inst[i].formula == F_MANDELBOX && !box) {
            sd.inst[i].p.p[0] = D6::var(s.inst[i].p.p[0], 4);
            sd.inst[i].p.p[1] = D6::var(s.inst[i].p.p[1], 5);
            box = true;
        }
    }
    V3<double> o = V3<double>::from(u->camera_pos);
    for (int i = 0; i < n; ++i) {
        V3<double> d = primary_dir<double>(*u, pix[2 * i], pix[2 * i + 1], (int)u->resolution[0], (int)u->resolution[1]);
        V3<double> p = o + d * t_in[i];
        V3<D6> q(D6::var(p.x, 0), D6::var(p.y, 1), D6::var(p.z, 2));
        D6 de = scene_de(sd, q);
        double den = de.d[0] * d.x + de.d[1] * d.y + de.d[2] * d.z - mc.eps_c;
        const int slot[6] = {3, 4, 5, 0, 1, 2};
        for (int k = 0; k < 6; ++k) g_out[6 * i + k] = den != 0 ? -de.d[slot[k]] / den : 0.0;
    }
    return 0;
}

// Individual helpers for closed-form and invariant checks.
//   0 box_fold(z3, L) -> z3                    1 sphere_fold(z3, dr, m2, F2) -> z3, dr
//   2 sphere_invert_offset(z3, dr, c3, R2)     3 lattice_wrap(p3, cell) -> p3
//   4 sort_desc_abs(z3) -> z3                  5 rotate_xyz_deg(z3, deg3) -> z3
//   6 bulb_power_step(z3, dr, add3, n) -> z3, dr
//   7..12 smin, smax, ssub, smin_lin, smin_nlin, smix (a, b, k) -> r
//   13 combine_de(mode, d, di, k) -> d, h      14 trap_distance(shape, z3, c3, box3, r, tube)
//   16 formula_finalize(f, z3, dr) -> d  (formula steps: ndo_formula_step)
//   17 formula_iterations(f, p16) -> n         18 trap run: mode, shape, radius, iters, zs(3*iters)
//   20 hash12w(sx, sy) -> u                    21 cosine_sample(n3, sx, sy) -> dir3
//   22 spectral round trip (hero, rgb3) -> rgb3 23 gamut_fit(rgb3) -> rgb3
//   24 fresnel_schlick(cos, ior) -> f          25 hero wavelengths(hero) -> lambda8
int ndo_helper(int op, const double* a, double* out) {
    using D3 = V3<double>;
    switch (op) {
        case 0: { D3 r = box_fold(D3(a[0], a[1], a[2]), a[3]); out[0] = r.x; out[1] = r.y; out[2] = r.z; return 0; }
        case 1: { auto r = sphere_fold(D3(a[0], a[1], a[2]), a[3], a[4], a[5]);
                  out[0] = r.z.x; out[1] = r.z.y; out[2] = r.z.z; out[3] = r.dr; return 0; }
        case 2: { auto r = sphere_invert_offset(D3(a[0], a[1], a[2]), a[3], D3(a[4], a[5], a[6]), a[7]);
                  out[0] = r.z.x; out[1] = r.z.y; out[2] = r.z.z; out[3] = r.dr; return 0; }
        case 3: { D3 r = lattice_wrap(D3(a[0], a[1], a[2]), a[3]); out[0] = r.x; out[1] = r.y; out[2] = r.z; return 0; }
        case 4: { D3 r = sort_desc_abs(D3(a[0], a[1], a[2])); out[0] = r.x; out[1] = r.y; out[2] = r.z; return 0; }
        case 5: { D3 r = rotate_xyz_deg(D3(a[0], a[1], a[2]), a[3], a[4], a[5]); out[0] = r.x; out[1] = r.y; out[2] = r.z; return 0; }
        case 6: { Carry<double> c{D3(a[0], a[1], a[2]), a[3]};
                  auto r = bulb_power_step(c, D3(a[4], a[5], a[6]), a[7]);
                  out[0] = r.z.x; out[1] = r.z.y; out[2] = r.z.z; out[3] = r.dr; return 0; }
        case 7: out[0] = smin(a[0], a[1], a[2]); return 0;
        case 8: out[0] = smax(a[0], a[1], a[2]); return 0;
        case 9: out[0] = ssub(a[0], a[1], a[2]); return 0;
        case 10: out[0] = smin_lin(a[0], a[1], a[2]); return 0;
        case 11: out[0] = smin_nlin(a[0], a[1], a[2]); return 0;
        case 12: out[0] = smix(a[0], a[1], a[2]); return 0;
        case 13: { double d, h; combine_de((int)a[0], a[1], a[2], a[3], &d, &h); out[0] = d; out[1] = h; return 0; }
        case 14: { TrapSpec t{}; t.shape = (int)a[0];
                   for (int k = 0; k < 3; ++k) { t.center[k] = a[4 + k]; t.box[k] = a[7 + k]; }
                   t.radius = a[10]; t.tube = a[11];
                   out[0] = trap_distance(t, D3(a[1], a[2], a[3])); return 0; }
        case 16: { Carry<double> c{D3(a[1], a[2], a[3]), a[4]}; out[0] = formula_finalize((int)a[0], c); return 0; }
        case 17: { Params<double> p; for (int k = 0; k < 16; ++k) p.p[k] = a[1 + k];
                   out[0] = formula_iterations((int)a[0], p); return 0; }
        case 18: { TrapSpec t{}; t.mode = (int)a[0]; t.shape = (int)a[1]; t.radius = a[2]; t.tube = 0;
                   t.box[0] = t.box[1] = t.box[2] = 1;
                   int iters = (int)a[3];
                   TrapAcc<double> acc = trap_begin<double>(t);
                   D3 prev(a[4], a[5], a[6]);
                   for (int it = 0; it < iters; ++it) {
                       D3 z(a[7 + 3 * it], a[8 + 3 * it], a[9 + 3 * it]);
                       acc = trap_update(acc, t, z, prev, it);
                       prev = z;
                   }
                   out[0] = trap_finish(acc, t, iters); return 0; }
        case 20: out[0] = hash12w((float)a[0], (float)a[1]); return 0;
        case 21: { D3 r = cosine_sample(normalize(D3(a[0], a[1], a[2])), Seed{(float)a[3], (float)a[4]});
                   out[0] = r.x; out[1] = r.y; out[2] = r.z; return 0; }
        case 22: { Spectral<double> sp(a[0]);
                   sp.to_rgb(sp.from_rgb(a[1], a[2], a[3]), out); return 0; }
        case 23: { out[0] = a[0]; out[1] = a[1]; out[2] = a[2]; gamut_fit(out); return 0; }
        case 24: out[0] = fresnel_schlick(a[0], a[1]); return 0;
        case 25: { Spectral<double> sp(a[0]); for (int k = 0; k < 8; ++k) out[k] = sp.lambda[k]; return 0; }
        case 26: out[0] = cauchy_ior(a[0], a[1], a[2]); return 0;
        case 27: { bool tir; D3 r = refract_safe(D3(a[0], a[1], a[2]), D3(a[3], a[4], a[5]), a[6], &tir);
                   out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = tir ? 1.0 : 0.0; return 0; }
        case 29: { Spectral<double> sp(a[0]); Spec<double> v = sp.from_rgb(a[1], a[2], a[3]);
                   for (int k = 0; k < 8; ++k) out[k] = v.v[k];
                   return 0; }
        case 30: { Spectral<double> sp(a[0]); Mat m = default_mat(); m.ior = a[2]; m.film_ior = a[3];
                   m.film_thickness = a[4]; m.film_strength = a[5]; m.film_angle_scale = a[6];
                   Spec<double> w = thin_film_tint(sp, a[1], m);
                   for (int k = 0; k < 8; ++k) out[k] = w.v[k];
                   return 0; }
        case 31: { D3 r = film_noise(D3(a[0], a[1], a[2])); out[0] = r.x; out[1] = r.y; out[2] = r.z; return 0; }
        case 32: { float h[3]; hash33(a[0], a[1], a[2], h); out[0] = h[0]; out[1] = h[1]; out[2] = h[2]; return 0; }
        case 33: { D3 r = oct_quantize(D3(a[0], a[1], a[2])); out[0] = r.x; out[1] = r.y; out[2] = r.z; return 0; }
        case 34: { Subsurface<double> ss{};  // diffuse_weight: amount, n.l, blurred n.l, 8 reaches
                   ss.amount = a[0]; ss.radius = 0.1; ss.blur_n = D3(1.0, 0.0, 0.0);
                   for (int k = 0; k < 8; ++k) ss.reach.v[k] = a[3 + k];
                   const double c = a[2], sn = std::sqrt(std::max(0.0, 1.0 - c * c));
                   Spec<double> w = diffuse_weight(ss, a[1], D3(c, sn, 0.0));
                   for (int k = 0; k < 8; ++k) out[k] = w.v[k];
                   return 0; }
        case 36: {  // light_terms: l, n, view, glossiness (spec power mix(4, 128, glossiness)) -> n.l, specular
                   const double g = a[9], power = 4.0 + (128.0 - 4.0) * std::min(std::max(g, 0.0), 1.0);
                   light_terms(normalize(D3(a[0], a[1], a[2])), normalize(D3(a[3], a[4], a[5])),
                               normalize(D3(a[6], a[7], a[8])), power, g, true, &out[0], &out[1]);
                   return 0; }
        case 37: { D3 r = perturb_normal(normalize(D3(a[0], a[1], a[2])), a[3], Seed{(float)a[4], (float)a[5]});
                   out[0] = r.x; out[1] = r.y; out[2] = r.z; return 0; }
        case 38: { double sc = 0; out[0] = roulette(a[0], Seed{(float)a[1], (float)a[2]}, &sc) ? 1.0 : 0.0; out[1] = sc;
                   return 0; }
        case 39: { D3 r = undo_tonemap(D3(a[0], a[1], a[2])); out[0] = r.x; out[1] = r.y; out[2] = r.z; return 0; }
        case 40: {  // light_sample: type, position_or_direction, colour, brightness, ray_direction, spread,
                    // waist, max_dist, surface point -> direction, colour, distance
                   nd_light L{};
                   L.light_type = (float)a[0];
                   for (int k = 0; k < 3; ++k) {
                       L.position_or_direction[k] = (float)a[1 + k];
                       L.color[k] = (float)a[4 + k];
                       L.ray_direction[k] = (float)a[8 + k];
                   }
                   L.brightness = (float)a[7];
                   L.spread = (float)a[11];
                   L.waist = (float)a[12];
                   D3 dir, col;
                   double dist = 0;
                   light_sample(L, a[13], D3(a[14], a[15], a[16]), &dir, &col, &dist);
                   out[0] = dir.x; out[1] = dir.y; out[2] = dir.z;
                   out[3] = col.x; out[4] = col.y; out[5] = col.z;
                   out[6] = dist;
                   return 0; }
        case 35: { Subsurface<double> ss{};  // subsurface_transmit: radius, depth, 8 reaches
                   ss.amount = 1.0; ss.radius = a[0];
                   for (int k = 0; k < 8; ++k) ss.reach.v[k] = a[2 + k];
                   Spec<double> w = subsurface_transmit(ss, a[1]);
                   for (int k = 0; k < 8; ++k) out[k] = w.v[k];
                   return 0; }
        case 28: { Spectral<double> sp(a[0]); Mat m = default_mat(); m.ior = a[7]; m.abbe = a[8];
                   Spec<double> w = dispersion_weights(sp, D3(a[1], a[2], a[3]), D3(a[4], a[5], a[6]), m, a[9] > 0.5,
                                                       D3(a[10], a[11], a[12]), a[13]);
                   for (int k = 0; k < 8; ++k) out[k] = w.v[k];
                   return 0; }
        // depth of field: dof_aperture(hit_t, focus_t) with aperture a[2], focus_range a[3]
        case 60: { nd_uniforms u{}; u.aperture = (float)a[2]; u.focus_range = (float)a[3];
                   out[0] = dof_aperture(u, a[0], a[1]); return 0; }
        // hash12 (IGN) and the lens disk sample of the float seed (a[0], a[1])
        case 61: { const float sx = (float)a[0], sy = (float)a[1];
                   const double r = std::sqrt((double)hash12(sx, sy));
                   const double th = (double)hash12w(sx + 17.17f, sy + 71.71f) * 6.28318530718;
                   out[0] = hash12(sx, sy); out[1] = r * std::cos(th); out[2] = r * std::sin(th); return 0; }
    }
    return -1;
}

// march_inside from origin along dir with a step budget: out = (exited, t, evals).
int ndo_march_inside(const nd_uniforms* u, const int32_t* cm, const double* origin, const double* dir, int steps,
                     double* out) {
    Scene<double> s = decode_scene<double>(*u, code_map(cm));
    int ev = 0;
    MarchResult<double> r = march_inside(s, V3<double>(origin[0], origin[1], origin[2]), V3<double>(dir[0], dir[1], dir[2]), steps,
                                         (double)u->max_dist, &ev);
    out[0] = r.hit ? 1.0 : 0.0; out[1] = r.dist; out[2] = ev;
    return 0;
}

// The path trace's spectral shadow visibility (8 lanes of the hero wavelength's basis).
int ndo_visibility(const nd_uniforms* u, const int32_t* cm, double hero, const double* origin, const double* dir,
                   double light_dist, double k, int steps, double* out) {
    Scene<double> s = decode_scene<double>(*u, code_map(cm));
    Spectral<double> sp(hero);
    Spec<double> v = visibility_march(*u, s, sp, V3<double>(origin[0], origin[1], origin[2]), V3<double>(dir[0], dir[1], dir[2]),
                                      light_dist, k, steps, (int*)nullptr);
    for (int i = 0; i < 8; ++i) out[i] = v.v[i];
    return 0;
}

// subsurface_light_path from pos (surface normal n) toward light `light`, for a material of
// colour albedo_rgb with the given subsurface amount and radius: out[0..7] = the spectral
// transmission, out[8] = DE evaluations, out[9..11] = the blurred normal, out[12..19] = reach.
int ndo_subsurface_path(const nd_uniforms* u, const int32_t* cm, double hero, const double* pos, const double* n,
                        int start_inside, int light, int occlusion, double amount, double radius,
                        const double* albedo_rgb, double* out) {
    Scene<double> s = decode_scene<double>(*u, code_map(cm));
    Spectral<double> sp(hero);
    Mat m = default_mat();
    for (int k = 0; k < 3; ++k) m.color[k] = albedo_rgb[k];
    m.subsurface = amount;
    m.subsurface_radius = radius;
    const V3<double> p(pos[0], pos[1], pos[2]), nn(n[0], n[1], n[2]);
    int ev = 0;
    const Subsurface<double> ss = subsurface_setup(s, p, nn, m, sp.from_rgb(m.color[0], m.color[1], m.color[2]), &ev);
    V3<double> ld, col;
    double dist;
    light_sample(u->lights[light], (double)u->max_dist, p, &ld, &col, &dist);
    Spec<double> v = subsurface_light_path(*u, s, sp, p, nn, start_inside != 0, ss, u->lights[light], ld, dist,
                                           occlusion != 0, &ev);
    for (int k = 0; k < 8; ++k) out[k] = v.v[k];
    out[8] = ev;
    out[9] = ss.blur_n.x; out[10] = ss.blur_n.y; out[11] = ss.blur_n.z;
    for (int k = 0; k < 8; ++k) out[12 + k] = ss.reach.v[k];
    return 0;
}

// placeCascades for a camera position (out: 8 x [corner xyz, cell]).
int ndo_place_cascades(const double* cam, double max_dist, int levels, int res, double* out) {
    double o[ND_MAX_CASCADES][4];
    place_cascades(cam, max_dist, levels, res, o);
    std::memcpy(out, o, sizeof(o));
    return 0;
}

// The accelerator grid the oracle builds for u (prec 0: FP64, 1: FP32), levels*res^3 values
// in [level][z][y][x] order, and accel_skip at n points.
int ndo_accel(const nd_uniforms* u, const int32_t* cm, int prec, double* grid_out, const double* pts, int n,
              double* skip_out) {
    auto go = [&](auto tag) {
        using T = decltype(tag);
        Scene<T> s = decode_scene<T>(*u, code_map(cm));
        auto g = build_accel(*u, s);
        if (!g) return -1;
        if (grid_out) std::memcpy(grid_out, g->safe.data(), sizeof(double) * g->safe.size());
        for (int i = 0; i < n; ++i) skip_out[i] = (double)val(accel_skip(g.get(), V3<T>(T(pts[3 * i]), T(pts[3 * i + 1]), T(pts[3 * i + 2]))));
        return 0;
    };
    return prec == 0 ? go(0.0) : go(0.0f);
}

// formula step with an explicit layout: a = [f, z3, dr, pos3, p16]
int ndo_formula_step(int f, const double* z, double dr, const double* pos, const double* p16, double* out) {
    Params<double> p;
    for (int k = 0; k < 16; ++k) p.p[k] = p16[k];
    Carry<double> c{V3<double>(z[0], z[1], z[2]), dr};
    auto r = formula_step(f, c, V3<double>(pos[0], pos[1], pos[2]), p);
    out[0] = r.z.x; out[1] = r.z.y; out[2] = r.z.z; out[3] = r.dr;
    return 0;
}

}  // extern "C"
