// nd_mesh.hip - the mesher's distance-field passes: cs_mesh_points (mesh.wgsl:49-61), the DE
// minus the iso level on blocks of a regular grid, and cs_mesh_probe (mesh.wgsl:63-92), the
// DE and its tetrahedral gradient at free points. The host side they replace is
// MeshGpu.run(.points / .probe) (src/gpu/mesher.zig:229-; export_mesh.zig:134-160,
// play_mode.zig:389-402).
//
// One thread per point. With span 8 a wave covers one z-slab of an 8x8x8 block (x fastest),
// so neighbouring lanes evaluate neighbouring points and the scene parameters are read
// through wave-uniform addresses.
#include <hip/hip_runtime.h>

#include <cstring>

#include "nd_ctx.h"
#include "nd_de.hpp"

namespace nd {

struct MeshGridDev {
    float ox, oy, oz, cell, iso;
    uint32_t span, local_offset, per_item;
};

template <bool EXT>
__global__ void __launch_bounds__(256) k_mesh_points(const DevScene* __restrict__ sc, const uint32_t* __restrict__ blocks,
                                                     MeshGridDev g, uint32_t total, float* __restrict__ out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint32_t item = t / g.per_item, l = t - item * g.per_item;
    const uint32_t lx = l % g.span, ly = (l / g.span) % g.span, lz = l / (g.span * g.span);
    // idx = block * MESH_BLOCK + local + local_offset; p = origin + idx * cell
    const uint32_t ix = blocks[4 * item] * 8u + lx + g.local_offset;
    const uint32_t iy = blocks[4 * item + 1] * 8u + ly + g.local_offset;
    const uint32_t iz = blocks[4 * item + 2] * 8u + lz + g.local_offset;
    const V3<float> p = mk3(add_rn(g.ox, mul_rn((float)ix, g.cell)), add_rn(g.oy, mul_rn((float)iy, g.cell)),
                            add_rn(g.oz, mul_rn((float)iz, g.cell)));
    out[t] = scene_de<false, EXT>(*sc, p, nullptr) - g.iso;
}

// (DE at p, tetrahedral gradient / (4 |h|)) per point (x, y, z, h); h < 0 turns the carves off
// for that point (carves_off = raw.w < 0). sc_bare is the same scene without carves.
template <bool EXT>
__global__ void __launch_bounds__(256) k_mesh_probe(const DevScene* __restrict__ sc, const DevScene* __restrict__ sc_bare,
                                                    const float* __restrict__ pts, uint32_t n, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float w = pts[4 * i + 3];
    const DevScene& s = w < 0.f ? *sc_bare : *sc;
    const float h = fabsf(w);
    const V3<float> q = mk3(pts[4 * i], pts[4 * i + 1], pts[4 * i + 2]);
    const float K[4][3] = {{1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {1, 1, 1}};
    const float centre = scene_de<false, EXT>(s, q, nullptr);
    V3<float> g = mk3(0.f, 0.f, 0.f);
    for (int k = 0; k < 4; ++k) {
        const float d = scene_de<false, EXT>(s, mk3(q.x + K[k][0] * h, q.y + K[k][1] * h, q.z + K[k][2] * h), nullptr);
        g = mk3(g.x + K[k][0] * d, g.y + K[k][1] * d, g.z + K[k][2] * d);
    }
    const float den = 4.f * h;  // TETRA_GRAD_DIVISOR
    out[4 * i] = centre;
    out[4 * i + 1] = g.x / den;
    out[4 * i + 2] = g.y / den;
    out[4 * i + 3] = g.z / den;
}

}  // namespace nd

namespace {
// The scene of the uniforms as the march passes see it (formula table, hoisted terms).
nd_status mesh_scene(nd_ctx* c, const void* uniforms, size_t size, DevScene* out) {
    if (!uniforms || size != sizeof(nd_uniforms)) return nd_fail(c, ND_E_INVALID, "uniforms_size %zu != %zu", size, sizeof(nd_uniforms));
    if (hipSetDevice(c->device) != hipSuccess) return nd_fail(c, ND_E_DEVICE, "hipSetDevice");
    nd_march_opts o{ND_NORMAL_TETRA, 0, 0, ND_KERNEL_PERSISTENT};
    DevJob job;
    std::memset(&job, 0, sizeof(job));
    nd_status st = nd_pack_job(c, *(const nd_uniforms*)uniforms, o, nd_rect{0, 0, 0, 0}, job);
    if (st != ND_OK) return st;
    *out = job.scene;
    return ND_OK;
}
// Device copies of host inputs and outputs for one call; freed by the destructor.
struct Scratch {
    void* p[4] = {nullptr, nullptr, nullptr, nullptr};
    int n = 0;
    hipError_t alloc(void** dst, size_t bytes) {
        hipError_t e = hipMalloc(dst, bytes);
        if (e == hipSuccess) p[n++] = *dst;
        return e;
    }
    ~Scratch() {
        for (int i = 0; i < n; ++i) (void)hipFree(p[i]);
    }
};
}  // namespace

extern "C" nd_status nd_mesh_points(nd_ctx* c, const void* uniforms, size_t uniforms_size, const nd_mesh_grid* grid,
                                    const uint32_t* blocks, uint32_t n_blocks, float* values, float* kernel_ms) {
    if (!c) return ND_E_INVALID;
    if (!grid) return nd_fail(c, ND_E_INVALID, "mesh grid is null");
    if (grid->span < 1 || grid->span > 8) return nd_fail(c, ND_E_INVALID, "mesh span %u (1..8)", grid->span);
    if (!(grid->cell > 0.f)) return nd_fail(c, ND_E_INVALID, "mesh cell %g", (double)grid->cell);
    if (n_blocks == 0) return ND_OK;
    if (!blocks || !values) return nd_fail(c, ND_E_INVALID, "mesh points need blocks and values");
    const uint64_t per = (uint64_t)grid->span * grid->span * grid->span;
    const uint64_t total = per * n_blocks;
    if (total > (1ull << 31)) return nd_fail(c, ND_E_INVALID, "%llu mesh points exceed 2^31", (unsigned long long)total);
    DevScene sc;
    nd_status st = mesh_scene(c, uniforms, uniforms_size, &sc);
    if (st != ND_OK) return st;
    Scratch tmp;
    DevScene* d_sc = nullptr;
    const uint32_t* d_blocks = blocks;
    float* d_vals = values;
    hipError_t e = tmp.alloc((void**)&d_sc, sizeof(DevScene));
    if (e == hipSuccess) e = hipMemcpyAsync(d_sc, &sc, sizeof(DevScene), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && !grid->device_ptrs) {
        uint32_t* b = nullptr;
        e = tmp.alloc((void**)&b, sizeof(uint32_t) * 4 * (size_t)n_blocks);
        if (e == hipSuccess) e = hipMemcpyAsync(b, blocks, sizeof(uint32_t) * 4 * (size_t)n_blocks, hipMemcpyHostToDevice, c->stream);
        d_blocks = b;
        if (e == hipSuccess) e = tmp.alloc((void**)&d_vals, sizeof(float) * total);
    }
    if (e != hipSuccess) return nd_hip_fail(c, e, "mesh points set-up");
    nd::MeshGridDev g{grid->origin[0], grid->origin[1], grid->origin[2], grid->cell, grid->iso, grid->span,
                      grid->local_offset, (uint32_t)per};
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (kernel_ms) {
        if ((e = hipEventCreate(&ev[0])) == hipSuccess) e = hipEventCreate(&ev[1]);
        if (e == hipSuccess) e = hipEventRecord(ev[0], c->stream);
    }
    const dim3 gr((uint32_t)((total + 255) / 256));
    if (e == hipSuccess) {
        if (nd_scene_ext(sc)) hipLaunchKernelGGL(nd::k_mesh_points<true>, gr, dim3(256), 0, c->stream, d_sc, d_blocks, g, (uint32_t)total, d_vals);
        else hipLaunchKernelGGL(nd::k_mesh_points<false>, gr, dim3(256), 0, c->stream, d_sc, d_blocks, g, (uint32_t)total, d_vals);
        e = hipGetLastError();
    }
    if (e == hipSuccess && kernel_ms) e = hipEventRecord(ev[1], c->stream);
    if (e == hipSuccess && !grid->device_ptrs)
        e = hipMemcpyAsync(values, d_vals, sizeof(float) * total, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && kernel_ms) e = hipEventElapsedTime(kernel_ms, ev[0], ev[1]);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (e != hipSuccess) return nd_hip_fail(c, e, "mesh points");
    return ND_OK;
}

extern "C" nd_status nd_mesh_probe(nd_ctx* c, const void* uniforms, size_t uniforms_size, const float* points, uint32_t n,
                                   float* out, uint32_t device_ptrs) {
    if (!c) return ND_E_INVALID;
    if (n == 0) return ND_OK;
    if (!points || !out) return nd_fail(c, ND_E_INVALID, "mesh probe needs points and an output");
    DevScene sc[2];
    nd_status st = mesh_scene(c, uniforms, uniforms_size, &sc[0]);
    if (st != ND_OK) return st;
    sc[1] = sc[0];
    sc[1].carve_count = 0;  // carves_off
    Scratch tmp;
    DevScene* d_sc = nullptr;
    const float* d_pts = points;
    float* d_out = out;
    hipError_t e = tmp.alloc((void**)&d_sc, 2 * sizeof(DevScene));
    if (e == hipSuccess) e = hipMemcpyAsync(d_sc, sc, 2 * sizeof(DevScene), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && !device_ptrs) {
        float* p = nullptr;
        e = tmp.alloc((void**)&p, sizeof(float) * 4 * (size_t)n);
        if (e == hipSuccess) e = hipMemcpyAsync(p, points, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice, c->stream);
        d_pts = p;
        if (e == hipSuccess) e = tmp.alloc((void**)&d_out, sizeof(float) * 4 * (size_t)n);
    }
    if (e != hipSuccess) return nd_hip_fail(c, e, "mesh probe set-up");
    const dim3 gr((n + 255) / 256);
    if (nd_scene_ext(sc[0])) hipLaunchKernelGGL(nd::k_mesh_probe<true>, gr, dim3(256), 0, c->stream, d_sc, d_sc + 1, d_pts, n, d_out);
    else hipLaunchKernelGGL(nd::k_mesh_probe<false>, gr, dim3(256), 0, c->stream, d_sc, d_sc + 1, d_pts, n, d_out);
    e = hipGetLastError();
    if (e == hipSuccess && !device_ptrs) e = hipMemcpyAsync(out, d_out, sizeof(float) * 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return nd_hip_fail(c, e, "mesh probe");
    return ND_OK;
}
