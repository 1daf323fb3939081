// nd_accel.hip - building the empty-space accelerator grids of a batch (cs_build_accel,
// template.wgsl:5409-5431; the host side of buildAccel, fr.zig:3172-3212).
//
// One grid per distinct (scene, cascades): jobs that share the scene and the camera-centred
// cascades share the grid. The grids of the last batch stay cached on the device and are
// reused while every job of a later batch finds its grid there.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "nd_ctx.h"
#include "nd_accel.hpp"
#include "nd_de.hpp"

namespace nd {

struct AccelDesc {
    float prm[ND_MAX_CASCADES][4];
    int res, levels;
    float safety;
};

// One thread per voxel, voxels in [level][z][y][x] order: the safe skip distance
// max(DE(centre) - HALF_SQRT3 cell, 0) * safety. The centre and the bound are formed with
// separate roundings (no fused multiply-add), as the scalar restatement forms them.
template <bool EXT>
__global__ void k_accel_build(const DevScene* __restrict__ sc, float* __restrict__ grid, AccelDesc d, uint32_t n) {
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const uint32_t res = (uint32_t)d.res;
    const uint32_t x = idx % res, y = (idx / res) % res, gz = idx / (res * res);
    const uint32_t level = gz / res, cz = gz % res;
    const float cell = d.prm[level][3];
    const V3<float> c = mk3(add_rn(d.prm[level][0], mul_rn((float)x + 0.5f, cell)),
                            add_rn(d.prm[level][1], mul_rn((float)y + 0.5f, cell)),
                            add_rn(d.prm[level][2], mul_rn((float)cz + 0.5f, cell)));
    const float de = scene_de<false, EXT>(*sc, c, nullptr);
    grid[idx] = mul_rn(fmaxf(sub_rn(de, mul_rn(0.8660254f, cell)), 0.f), d.safety);
}

}  // namespace nd

static uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) { h ^= b[i]; h *= 1099511628211ull; }
    return h;
}

static uint64_t accel_key(const DevJob& j) {
    uint64_t h = 1469598103934665603ull;
    h = fnv1a(h, &j.scene, sizeof(DevScene));
    h = fnv1a(h, &j.accel_res, sizeof(int));
    h = fnv1a(h, &j.accel_levels, sizeof(int));
    h = fnv1a(h, &j.accel_safety, sizeof(float));
    h = fnv1a(h, j.accel_prm, sizeof(j.accel_prm));
    return h;
}

nd_status nd_accel_prepare(nd_ctx* c, DevJob* jobs, uint32_t n) {
    std::vector<uint64_t> keys(n, 0);
    std::vector<uint32_t> uniq_first;
    std::vector<uint64_t> uniq;
    for (uint32_t i = 0; i < n; ++i) {
        jobs[i].accel = nullptr;
        if (jobs[i].accel_levels <= 0) continue;
        keys[i] = accel_key(jobs[i]);
        bool seen = false;
        for (uint64_t k : uniq) seen = seen || k == keys[i];
        if (!seen) { uniq.push_back(keys[i]); uniq_first.push_back(i); }
    }
    if (uniq.empty()) return ND_OK;
    auto find = [&](uint64_t k) -> const NdAccelEntry* {
        for (const NdAccelEntry& e : c->accel_cache)
            if (e.key == k) return &e;
        return nullptr;
    };
    bool cached = true;
    for (uint64_t k : uniq) cached = cached && find(k) != nullptr;
    if (!cached) {
        size_t total = 0;
        for (uint32_t f : uniq_first) {
            const size_t r = (size_t)jobs[f].accel_res;
            total += r * r * r * (size_t)jobs[f].accel_levels;
        }
        if (total > ((size_t)1 << 31)) return nd_fail(c, ND_E_INVALID, "accelerator grids of %zu cells exceed 2^31", total);
        c->accel_cache.clear();
        nd_status st = c->accel.ensure(c, total * sizeof(float));
        if (st != ND_OK) return st;
        if ((st = c->accel_scenes.ensure(c, uniq.size() * sizeof(DevScene))) != ND_OK) return st;
        c->accel_scene_host.resize(uniq.size());
        for (size_t u = 0; u < uniq.size(); ++u) c->accel_scene_host[u] = jobs[uniq_first[u]].scene;
        hipError_t e = hipMemcpyAsync(c->accel_scenes.p, c->accel_scene_host.data(), uniq.size() * sizeof(DevScene),
                                      hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return nd_hip_fail(c, e, "upload accelerator scenes");
        size_t off = 0;
        for (size_t u = 0; u < uniq.size(); ++u) {
            const DevJob& j = jobs[uniq_first[u]];
            nd::AccelDesc d;
            std::memcpy(d.prm, j.accel_prm, sizeof(d.prm));
            d.res = j.accel_res;
            d.levels = j.accel_levels;
            d.safety = j.accel_safety;
            const size_t cnt = (size_t)d.res * d.res * d.res * d.levels;
            const dim3 g((uint32_t)((cnt + 255) / 256));
            const DevScene* dsc = (const DevScene*)c->accel_scenes.p + u;
            if (nd_scene_ext(j.scene))
                hipLaunchKernelGGL(nd::k_accel_build<true>, g, dim3(256), 0, c->stream, dsc, (float*)c->accel.p + off, d, (uint32_t)cnt);
            else
                hipLaunchKernelGGL(nd::k_accel_build<false>, g, dim3(256), 0, c->stream, dsc, (float*)c->accel.p + off, d, (uint32_t)cnt);
            if ((e = hipGetLastError()) != hipSuccess) return nd_hip_fail(c, e, "launch accelerator build");
            c->accel_cache.push_back(NdAccelEntry{uniq[u], off, cnt});
            off += cnt;
        }
    }
    for (uint32_t i = 0; i < n; ++i) {
        if (jobs[i].accel_levels <= 0) continue;
        jobs[i].accel = (const float*)c->accel.p + find(keys[i])->offset;
    }
    return ND_OK;
}

namespace nd {
// accel_skip at n points (xyz triples) with the grid of job J, read the way a ray reads it:
// through the register cache of the cell of the previous point (point i - 1), so points in
// ray order exercise the cache's box test and shuffled points the full lookup.
__global__ void k_accel_probe(const DevJob* __restrict__ J, const float* __restrict__ pts, uint32_t n,
                              float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const AccelView av = accel_view(*J);
    if (!av.grid) { out[i] = 0.f; return; }
    AccelCache c;
    accel_cache_reset(c);
    if (i > 0) (void)accel_skip_cached(av, c, mk3(pts[3 * i - 3], pts[3 * i - 2], pts[3 * i - 1]));
    out[i] = accel_skip_cached(av, c, mk3(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]));
}
}  // namespace nd

extern "C" nd_status nd_accel_probe(nd_ctx* c, const void* uniforms, size_t size, const float* pts, uint32_t n,
                                    float* skip_out) {
    if (!c) return ND_E_INVALID;
    if (!uniforms || size != sizeof(nd_uniforms)) return nd_fail(c, ND_E_INVALID, "uniforms_size %zu != %zu", size, sizeof(nd_uniforms));
    if (n == 0) return ND_OK;
    if (!pts || !skip_out) return nd_fail(c, ND_E_INVALID, "accel probe needs points and an output");
    if (hipSetDevice(c->device) != hipSuccess) return nd_fail(c, ND_E_DEVICE, "hipSetDevice");
    const nd_uniforms& u = *(const nd_uniforms*)uniforms;
    if (u.accel_enabled < 0.5f) return nd_fail(c, ND_E_INVALID, "the accelerator is off in these uniforms");
    nd_march_opts o{ND_NORMAL_TETRA, 0, 0, ND_KERNEL_PERSISTENT};
    DevJob job;
    std::memset(&job, 0, sizeof(job));
    nd_status st = nd_pack_job(c, u, o, nd_rect{0, 0, 0, 0}, job);
    if (st != ND_OK) return st;
    if ((st = nd_accel_prepare(c, &job, 1)) != ND_OK) return st;
    DevJob* d_job = nullptr;
    float *d_pts = nullptr, *d_out = nullptr;
    hipError_t e = hipMalloc((void**)&d_job, sizeof(DevJob));
    if (e == hipSuccess) e = hipMalloc((void**)&d_pts, sizeof(float) * 3 * (size_t)n);
    if (e == hipSuccess) e = hipMalloc((void**)&d_out, sizeof(float) * (size_t)n);
    if (e == hipSuccess) e = hipMemcpyAsync(d_job, &job, sizeof(DevJob), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_pts, pts, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(nd::k_accel_probe, dim3((n + 255) / 256), dim3(256), 0, c->stream, d_job, d_pts, n, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(skip_out, d_out, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_job);
    (void)hipFree(d_pts);
    (void)hipFree(d_out);
    if (e != hipSuccess) return nd_hip_fail(c, e, "accelerator probe");
    return ND_OK;
}
