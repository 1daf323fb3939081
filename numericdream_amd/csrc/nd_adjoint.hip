// nd_adjoint.hip - gradient of hit depth with respect to formula parameters and camera
// position, without storing march steps.
//
// At a hit the march has stopped where DE(o + t d; theta) = eps(t) = c t + f (the on-hit
// bisection of march_to, template.wgsl:1924-1931). The implicit-function theorem on
// F(t, theta) = DE(o + t d; theta) - c t - f gives
//     dt/dtheta = -(dDE/dtheta) / (grad DE . d - c),
//     dt/do     = -(grad DE)    / (grad DE . d - c)       (camera translation),
// with dDE/dtheta and grad DE from one forward-mode evaluation of the DE with six tangents
// (x, y, z, Mandelbulb power, Mandelbox scale, Mandelbox min radius). The loss is
// L = sum over pixels hit in both renders of (t - t_target)^2; each wave reduces its part in
// FP64 and adds it with one atomic per component.
//
// nd_radiance_adjoint (at the end) does the same for the direct-lighting colour; its kernel
// lives with the path trace's shading (nd_mc.hip).
//
// There is no counterpart in the reference (it renders, it does not differentiate).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "nd_ctx.h"
#include "nd_de.hpp"
#include "nd_internal.h"
#include "nd_rays.hpp"

namespace nd {

struct AdjArgs {
    const DevJob* jobs;
    const DevScene* shared;
    const DevCounters* ctr;
    DevHitBuffers hb;
    const float* target;       // all views, view v at pix_base[v]
    float* dtdp;               // 6 floats per pixel, or null
    const uint64_t* pix_base;  // first pixel of each view
    double* acc;               // [loss, 6 gradient components, pixels used]
};

template <bool SHARED, bool EXT>
__global__ void __launch_bounds__(64) k_adjoint(AdjArgs a, const DevJob* __restrict__ jobs,
                                                 const DevScene* __restrict__ shared) {
    using D6 = DualN<6>;
    const uint32_t n_hits = a.ctr->hit_count;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t i = blockIdx.x * 64 + threadIdx.x; i - threadIdx.x < n_hits; i += gridDim.x * 64) {
        if (i >= n_hits) continue;
        const uint32_t r = a.hb.ray[i];
        if (r == ~0u) continue;  // empty slot of a partly used hit chunk
        const int job = a.hb.job[i];
        const DevJob& J = jobs[job];
        const DevScene& sc = SHARED ? *shared : J.scene;
        int x, y;
        slot_pixel(J, r - J.ray_base, &x, &y);
        const size_t o = (size_t)y * J.rect_w + x;
        const V3<float> dir = primary_dir(J, J.rect_x + x, J.rect_y + y);
        const float t = a.hb.t[i];
        const V3<float> hp = mk3(J.cam[0] + dir.x * t, J.cam[1] + dir.y * t, J.cam[2] + dir.z * t);
        V3<D6> q = mk3(dvar<6>(hp.x, 0), dvar<6>(hp.y, 1), dvar<6>(hp.z, 2));
        const D6 d = scene_de_d<D6, EXT>(sc, q);
        const float den = d.d[0] * dir.x + d.d[1] * dir.y + d.d[2] * dir.z - J.eps_c;
        float g[6] = {0, 0, 0, 0, 0, 0};
        if (den != 0.f) {
            const float inv = -1.f / den;
            g[0] = d.d[3] * inv;  // Mandelbulb power
            g[1] = d.d[4] * inv;  // Mandelbox scale
            g[2] = d.d[5] * inv;  // Mandelbox min radius
            g[3] = d.d[0] * inv;  // camera x, y, z
            g[4] = d.d[1] * inv;
            g[5] = d.d[2] * inv;
        }
        const uint64_t pix = a.pix_base[job] + o;
        if (a.dtdp) {
#pragma unroll
            for (int k = 0; k < 6; ++k) a.dtdp[pix * 6 + k] = g[k];
        }
        const float tgt = a.target ? a.target[pix] : 0.f;
        if (tgt > 0.f && den != 0.f) {
            const double res = (double)t - (double)tgt;
            acc[0] += res * res;
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[1 + k] += 2.0 * res * (double)g[k];
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

__global__ void k_fill_f32(float* p, uint64_t n, float v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f64(double* p, int n) {
    int i = threadIdx.x;
    if (i < n) p[i] = 0.0;
}

}  // namespace nd

void nd_resident_adjoint(uint32_t* r) {
    int b = 0;
    r[RES_ADJOINT] = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, nd::k_adjoint<true, false>, 64, 0) == hipSuccess && b > 0 ? b : 8;
    b = 0;
    r[RES_ADJOINT_MIXED] =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, nd::k_adjoint<false, false>, 64, 0) == hipSuccess && b > 0 ? b : 8;
    b = 0;
    r[RES_ADJOINT_EXT] =
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, nd::k_adjoint<false, true>, 64, 0) == hipSuccess && b > 0 ? b : 8;
}

extern "C" nd_status nd_adjoint(nd_ctx* c, const nd_uniforms* views, uint32_t n, const nd_adjoint_io* io,
                                nd_batch_stats* stats) {
    if (!c) return ND_E_INVALID;
    if (!views || n == 0 || !io) return nd_fail(c, ND_E_INVALID, "adjoint needs views and an io record");
    if (n > 65535) return nd_fail(c, ND_E_INVALID, "at most 65535 views");
    hipError_t e;
    // pixel layout: view v occupies W_v * H_v consecutive pixels
    std::vector<uint64_t> base(n);
    uint64_t total = 0;
    for (uint32_t v = 0; v < n; ++v) {
        base[v] = total;
        uint64_t w = (uint64_t)views[v].resolution[0], h = (uint64_t)views[v].resolution[1];
        if (w == 0 || h == 0) return nd_fail(c, ND_E_INVALID, "view %u has an empty resolution", v);
        total += w * h;
    }
    // device work area: depth, target, dtdp, pixel bases, accumulators
    const bool dev = io->device_ptrs != 0;
    size_t need = 256 + 8 * 8 + 8 * (size_t)n + 256;
    if (!dev || !io->depth_out) need += 4 * total + 256;
    if (io->target_depth && !dev) need += 4 * total + 256;
    if (io->dtdp_out && !dev) need += 24 * total + 256;
    nd_status st = c->adj.ensure(c, need);
    if (st != ND_OK) return st;
    char* p = (char*)c->adj.p;
    auto carve = [&](size_t b) { char* r = p; p += (b + 255) & ~(size_t)255; return r; };
    double* acc = (double*)carve(8 * 8);
    uint64_t* d_base = (uint64_t*)carve(8 * (size_t)n);
    float* d_depth = (dev && io->depth_out) ? io->depth_out : (float*)carve(4 * total);
    const float* d_target = nullptr;
    if (io->target_depth) {
        if (dev) d_target = io->target_depth;
        else {
            float* t = (float*)carve(4 * total);
            if ((e = hipMemcpyAsync(t, io->target_depth, 4 * total, hipMemcpyHostToDevice, c->stream)) != hipSuccess)
                return nd_hip_fail(c, e, "upload target depth");
            d_target = t;
        }
    }
    float* d_dtdp = nullptr;
    if (io->dtdp_out) d_dtdp = dev ? io->dtdp_out : (float*)carve(24 * total);
    if ((e = hipMemcpyAsync(d_base, base.data(), 8 * (size_t)n, hipMemcpyHostToDevice, c->stream)) != hipSuccess)
        return nd_hip_fail(c, e, "upload view table");
    hipLaunchKernelGGL(nd::k_fill_f64, dim3(1), dim3(64), 0, c->stream, acc, 8);
    if (d_dtdp) hipLaunchKernelGGL(nd::k_fill_f32, dim3(1024), dim3(256), 0, c->stream, d_dtdp, 6 * total, 0.0f);

    // forward: the persistent march over all views (depth only)
    std::vector<nd_job> jobs(n);
    for (uint32_t v = 0; v < n; ++v) {
        jobs[v].uniforms = &views[v];
        jobs[v].rect = nd_rect{0, 0, 0, 0};
        std::memset(&jobs[v].out, 0, sizeof(nd_gbuffer_out));
        jobs[v].out.depth = d_depth + base[v];
        jobs[v].out.device_ptrs = 1;
    }
    nd_march_opts o{ND_NORMAL_ANALYTIC, 0, 0, ND_KERNEL_PERSISTENT};
    NdBatchExtra ex;
    ex.skip_surface = true;
    ex.adjoint_slots = true;
    nd_batch_stats fs{};
    if ((st = nd_batch_impl(c, jobs.data(), n, &o, &fs, &ex)) != ND_OK) return st;

    // adjoint over the compacted hits of that batch
    nd::AdjArgs a;
    a.jobs = (const DevJob*)c->jobs.p;
    a.shared = c->last_shared ? (const DevScene*)c->shared.p : nullptr;
    a.ctr = (const DevCounters*)c->ctr.p;
    a.hb = c->last_hb;  // the compacted hits of the batch just run
    a.target = d_target;
    a.dtdp = d_dtdp;
    a.pix_base = d_base;
    a.acc = acc;
    (void)hipEventRecord(c->ev[4], c->stream);
    // grid-stride over the hits, at most one per ray of the batch
    const int res = c->last_ext ? RES_ADJOINT_EXT : a.shared ? RES_ADJOINT : RES_ADJOINT_MIXED;
    const uint32_t grid = nd_pass_grid(c, res, ((uint64_t)c->last_rays + 63) / 64);
    if (c->last_ext) hipLaunchKernelGGL((nd::k_adjoint<false, true>), dim3(grid), dim3(64), 0, c->stream, a, a.jobs, a.shared);
    else if (a.shared) hipLaunchKernelGGL((nd::k_adjoint<true, false>), dim3(grid), dim3(64), 0, c->stream, a, a.jobs, a.shared);
    else hipLaunchKernelGGL((nd::k_adjoint<false, false>), dim3(grid), dim3(64), 0, c->stream, a, a.jobs, a.shared);
    if ((e = hipGetLastError()) != hipSuccess) return nd_hip_fail(c, e, "launch adjoint");
    (void)hipEventRecord(c->ev[5], c->stream);
    double h_acc[8];
    if ((e = hipMemcpyAsync(h_acc, acc, sizeof(h_acc), hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
        return nd_hip_fail(c, e, "read gradient");
    if (!dev) {
        if (io->depth_out && (e = hipMemcpyAsync(io->depth_out, d_depth, 4 * total, hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
            return nd_hip_fail(c, e, "read depth");
        if (io->dtdp_out && (e = hipMemcpyAsync(io->dtdp_out, d_dtdp, 24 * total, hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
            return nd_hip_fail(c, e, "read dt/dp");
    }
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return nd_hip_fail(c, e, "adjoint execution");
    float ms_adj = 0.f;
    (void)hipEventElapsedTime(&ms_adj, c->ev[4], c->ev[5]);
    if (io->loss_out) *io->loss_out = h_acc[0];
    if (io->grad_out) for (int k = 0; k < 6; ++k) io->grad_out[k] = h_acc[1 + k];
    if (io->pixels_used) *io->pixels_used = (uint64_t)h_acc[7];
    if (stats) {
        *stats = fs;
        stats->ms_surface = ms_adj;  // the adjoint pass takes the surface pass's slot
        stats->ms_total = fs.ms_total + ms_adj;
        stats->de_evals_surface = fs.hits;
        stats->de_evals += fs.hits;
        stats->kernels += 1;
    }
    return ND_OK;
}

// ----------------------------------------------------------------------------- radiance adjoint
// The scenes the radiance adjoint differentiates exactly: the material does not vary over the
// surface, and nothing but the local shading reaches the pixel.
static nd_status radiance_scope(nd_ctx* c, const nd_uniforms& u, uint32_t v) {
    const uint32_t parts = (uint32_t)u.debug_parts;
    if (u.mc_enabled > 0.5f) return nd_fail(c, ND_E_INVALID, "view %u: the radiance adjoint needs mc_enabled = 0", v);
    if (!(parts & ND_PART_REFLECTIONS))
        return nd_fail(c, ND_E_INVALID, "view %u: the radiance adjoint needs the reflection part off (PART_REFLECTIONS)", v);
    if (u.accel_enabled > 0.5f) return nd_fail(c, ND_E_UNSUPPORTED, "view %u: the radiance adjoint runs without the accelerator", v);
    if (u.dof_enabled > 0.5f && u.aperture > 0.0001f)
        return nd_fail(c, ND_E_UNSUPPORTED, "view %u: the radiance adjoint is the pinhole colour (depth of field off)", v);
    const nd_color_stop* first = nullptr;
    int first_count = -1;
    const int ni = std::min(std::max((int)u.instance_count, 0), ND_MAX_INSTANCES);
    for (int i = 0; i < ni; ++i) {
        const nd_fractal_instance& in = u.instances[i];
        if (in.hidden > 0.5f) continue;
        const int cc = (int)in.color_count;
        if (cc > 1) return nd_fail(c, ND_E_UNSUPPORTED, "view %u instance %d: colour strips vary the material over the surface", v, i);
        if (first_count < 0) { first_count = cc; first = cc ? &in.colors[0] : nullptr; }
        else if (cc != first_count || (cc && std::memcmp(first, &in.colors[0], sizeof(nd_color_stop)) != 0))
            return nd_fail(c, ND_E_UNSUPPORTED, "view %u: instances of different materials", v);
        if (cc) {
            const nd_color_stop& st = in.colors[0];
            if (st.roughness > 0.01f) return nd_fail(c, ND_E_UNSUPPORTED, "view %u instance %d: roughness above 0.01", v, i);
            if (st.transparency > 0.f && !(parts & ND_PART_REFRACTION))
                return nd_fail(c, ND_E_UNSUPPORTED, "view %u instance %d: transparency", v, i);
            if (st.film_strength > 0.001f && !(parts & ND_PART_IRIDESCENCE))
                return nd_fail(c, ND_E_UNSUPPORTED, "view %u instance %d: thin film", v, i);
            if (st.subsurface > 0.001f && !(parts & ND_PART_SUBSURFACE))
                return nd_fail(c, ND_E_UNSUPPORTED, "view %u instance %d: subsurface", v, i);
        }
    }
    return ND_OK;
}

extern "C" nd_status nd_radiance_adjoint(nd_ctx* c, const nd_uniforms* views, uint32_t n, const nd_radiance_io* io,
                                         nd_batch_stats* stats) {
    if (!c) return ND_E_INVALID;
    if (!views || n == 0 || !io) return nd_fail(c, ND_E_INVALID, "radiance adjoint needs views and an io record");
    if (n > 65535) return nd_fail(c, ND_E_INVALID, "at most 65535 views");
    nd_status st;
    for (uint32_t v = 0; v < n; ++v)
        if ((st = radiance_scope(c, views[v], v)) != ND_OK) return st;
    hipError_t e;
    std::vector<uint64_t> base(n);
    uint64_t total = 0;
    for (uint32_t v = 0; v < n; ++v) {
        base[v] = total;
        uint64_t w = (uint64_t)views[v].resolution[0], h = (uint64_t)views[v].resolution[1];
        if (w == 0 || h == 0) return nd_fail(c, ND_E_INVALID, "view %u has an empty resolution", v);
        total += w * h;
    }
    const bool dev = io->device_ptrs != 0;
    const size_t mcs = nd_mc_scene_bytes();
    size_t need = 256 + 8 * 8 + 8 * (size_t)n + 256 + mcs * n + 256 + 4 * total + 256;
    if (io->target_rgb && !dev) need += 12 * total + 256;
    if (io->rgb_out && !dev) need += 12 * total + 256;
    if (io->drgb_out && !dev) need += 72 * total + 256;
    if ((st = c->adj.ensure(c, need)) != ND_OK) return st;
    char* p = (char*)c->adj.p;
    auto carve = [&](size_t b) { char* r = p; p += (b + 255) & ~(size_t)255; return r; };
    double* acc = (double*)carve(8 * 8);
    uint64_t* d_base = (uint64_t*)carve(8 * (size_t)n);
    char* d_scenes = carve(mcs * n);
    float* d_depth = (float*)carve(4 * total);
    const float* d_target = nullptr;
    if (io->target_rgb) {
        if (dev) d_target = io->target_rgb;
        else {
            float* t = (float*)carve(12 * total);
            if ((e = hipMemcpyAsync(t, io->target_rgb, 12 * total, hipMemcpyHostToDevice, c->stream)) != hipSuccess)
                return nd_hip_fail(c, e, "upload target colour");
            d_target = t;
        }
    }
    float* d_rgb = io->rgb_out ? (dev ? io->rgb_out : (float*)carve(12 * total)) : nullptr;
    float* d_drgb = io->drgb_out ? (dev ? io->drgb_out : (float*)carve(72 * total)) : nullptr;
    std::vector<unsigned char> hs(mcs * n);
    for (uint32_t v = 0; v < n; ++v) nd_mc_scene_pack(views[v], hs.data() + mcs * v);
    if ((e = hipMemcpyAsync(d_base, base.data(), 8 * (size_t)n, hipMemcpyHostToDevice, c->stream)) != hipSuccess)
        return nd_hip_fail(c, e, "upload view table");
    if ((e = hipMemcpyAsync(d_scenes, hs.data(), mcs * n, hipMemcpyHostToDevice, c->stream)) != hipSuccess)
        return nd_hip_fail(c, e, "upload material scenes");
    hipLaunchKernelGGL(nd::k_fill_f64, dim3(1), dim3(64), 0, c->stream, acc, 8);
    if (d_rgb) hipLaunchKernelGGL(nd::k_fill_f32, dim3(1024), dim3(256), 0, c->stream, d_rgb, 3 * total, 0.0f);
    if (d_drgb) hipLaunchKernelGGL(nd::k_fill_f32, dim3(1024), dim3(256), 0, c->stream, d_drgb, 18 * total, 0.0f);

    // forward: the persistent march over all views (the hit list), parameters as adjoint slots
    std::vector<nd_job> jobs(n);
    for (uint32_t v = 0; v < n; ++v) {
        jobs[v].uniforms = &views[v];
        jobs[v].rect = nd_rect{0, 0, 0, 0};
        std::memset(&jobs[v].out, 0, sizeof(nd_gbuffer_out));
        jobs[v].out.depth = d_depth + base[v];
        jobs[v].out.device_ptrs = 1;
    }
    nd_march_opts o{ND_NORMAL_ANALYTIC, 0, 0, ND_KERNEL_PERSISTENT};
    NdBatchExtra ex;
    ex.skip_surface = true;
    ex.adjoint_slots = true;
    nd_batch_stats fs{};
    if ((st = nd_batch_impl(c, jobs.data(), n, &o, &fs, &ex)) != ND_OK) return st;

    NdRadianceLaunch L;
    L.jobs = (const DevJob*)c->jobs.p;
    L.shared = c->last_shared ? (const DevScene*)c->shared.p : nullptr;
    L.ctr = (const DevCounters*)c->ctr.p;
    L.hb = c->last_hb;
    L.scenes = d_scenes;
    L.target = d_target;
    L.rgb = d_rgb;
    L.drgb = d_drgb;
    L.pix_base = d_base;
    L.acc = acc;
    (void)hipEventRecord(c->ev[4], c->stream);
    const uint32_t res = nd_resident_radiance(c->last_ext);
    const uint64_t waves = ((uint64_t)c->last_rays + 63) / 64;
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(waves, (uint64_t)c->cu_count * res));
    if ((e = nd_launch_radiance_adjoint(L, grid, c->last_ext, c->stream)) != hipSuccess)
        return nd_hip_fail(c, e, "launch radiance adjoint");
    (void)hipEventRecord(c->ev[5], c->stream);
    double h_acc[8];
    if ((e = hipMemcpyAsync(h_acc, acc, sizeof(h_acc), hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
        return nd_hip_fail(c, e, "read gradient");
    if (!dev) {
        if (io->rgb_out && (e = hipMemcpyAsync(io->rgb_out, d_rgb, 12 * total, hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
            return nd_hip_fail(c, e, "read colour");
        if (io->drgb_out && (e = hipMemcpyAsync(io->drgb_out, d_drgb, 72 * total, hipMemcpyDeviceToHost, c->stream)) != hipSuccess)
            return nd_hip_fail(c, e, "read d colour");
    }
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return nd_hip_fail(c, e, "radiance adjoint execution");
    float ms_adj = 0.f;
    if (hipEventElapsedTime(&ms_adj, c->ev[4], c->ev[5]) != hipSuccess) { ms_adj = 0.f; (void)hipGetLastError(); }
    if (io->loss_out) *io->loss_out = h_acc[0];
    if (io->grad_out) for (int k = 0; k < 6; ++k) io->grad_out[k] = h_acc[1 + k];
    if (io->pixels_used) *io->pixels_used = (uint64_t)h_acc[7];
    if (stats) {
        *stats = fs;
        stats->ms_surface = ms_adj;  // the adjoint pass takes the surface pass's slot
        stats->ms_total = fs.ms_total + ms_adj;
        stats->kernels += 1;
    }
    return ND_OK;
}
