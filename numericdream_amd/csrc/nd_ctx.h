// nd_ctx.h - the context behind the opaque nd_ctx handle of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "nd_dev.h"
#include "nd_hip.h"
#include "nd_internal.h"

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    nd_status ensure(nd_ctx* c, size_t bytes);  // grows; a reallocation destroys c's captured graphs
    void release();
};

// Destroys the context's captured graphs (G-buffer passes and path-trace chunk passes).
void nd_drop_graphs(nd_ctx* c);

// A cached accelerator grid: its key (scene + cascades) and place in nd_ctx::accel.
struct NdAccelEntry {
    uint64_t key;
    size_t offset;  // in floats
    size_t count;
};

// A height-map image bound to a slot: its packed texel words (nd_image.cpp).
struct NdImage {
    std::vector<uint32_t> words;
    nd_image_desc desc{};
};

struct nd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    char err[512] = {0};  // the error sink (webgpu_context.zig:131-146 keeps 512 bytes too)
    int code_map[ND_MAX_CODES];
    nd_tuning tuning{};
    int cu_count = 256;
    size_t mem_total = 0;  // device memory in bytes (hipDeviceProp_t::totalGlobalMem)
    uint32_t resident[RES_COUNT] = {};  // resident waves per CU of each persistent kernel
    char arch[64] = {0};
    int wall_khz = 100000;
    uint32_t round_cap = 0xffffffffu;
    const volatile int* cancel = nullptr;
    char* noalloc_scratch = nullptr;   // this device's no-allocation scratch array (device address)
    hipEvent_t ev[8] = {};
    hipGraphExec_t graph_exec = nullptr;  // the persistent G-buffer passes of the last batch shape
                                          // (destroyed, with the path trace's, when a device
                                          // buffer is reallocated: both hold the old pointers)
    NdPassLaunch graph_key{};             // what that graph baked in
    NdPassLaunch graph_seen{};            // the launch record of the last eager batch: a repeat is captured
    bool graph_broken = false;            // capture failed once: batches launch pass by pass
    hipGraphExec_t mc_graph_exec = nullptr;         // the path trace's per-chunk passes
    alignas(8) unsigned char mc_graph_key[1024] = {};  // what that graph baked in
    DevBuf jobs, job_base, ctr, hits, shared, staging, latency, adj, mc;
    DevBuf color_scenes;             // per-job material scenes of the colour pass
    DevBuf mc_adapt;                 // adaptive sampling state per pixel of the last path-trace rect
    uint32_t mc_adapt_key[6] = {};   // that rect and target size (the state carries over to the next call)
    DevBuf accel, accel_scenes;              // accelerator grids of the last batch, their scenes
    std::vector<NdAccelEntry> accel_cache;
    std::vector<DevScene> accel_scene_host;
    bool last_shared = false;       // scene of the last batch was shared by all jobs
    bool last_ext = false;          // the last batch ran the kernels with every formula
    uint32_t last_rays = 0;         // rays of the last persistent batch
    DevHitBuffers last_hb{};        // hit list of the last persistent batch
    std::vector<DevJob> host_jobs;
    NdImage images[ND_IMAGE_SLOTS];  // bound height-map images
    DevBuf image_buf;                // their device layout: 16 slot headers of 8 words, then texels
    std::vector<uint32_t> image_words;  // host copy of image_buf
    std::vector<float> job_latency_ms;
};

// A fixed device region a batch carves all its buffers from instead of growing the context's
// buffers (nd_render_gbuffer_noalloc): nothing is allocated, sizes beyond it are rejected.
struct NdFixedWorkspace {
    char* base;          // device address (from hipGetSymbolAddress), 256-byte aligned
    size_t bytes;
    uint32_t max_rays;   // ray slots (64 per 8x8 tile) of the batch
    uint32_t max_grid;   // persistent workgroups per pass
};

// Internal options of a batch beyond the public nd_march_opts.
struct NdBatchExtra {
    bool skip_surface = false;   // no normal / trap pass (the adjoint needs only t)
    bool adjoint_slots = false;  // attach adjoint tangents to the differentiable params
    const NdFixedWorkspace* fixed = nullptr;
};

// Device address of the library's no-allocation scratch array on `device` (the current
// device), looked up once per device; nullptr if the lookup failed.
char* nd_noalloc_scratch(int device);
size_t nd_noalloc_scratch_bytes();

nd_status nd_batch_impl(nd_ctx* c, const nd_job* jobs, uint32_t n, const nd_march_opts* opts, nd_batch_stats* stats,
                        const NdBatchExtra* extra = nullptr);
nd_status nd_fail(nd_ctx* c, nd_status st, const char* fmt, ...);
nd_status nd_hip_fail(nd_ctx* c, hipError_t e, const char* what);
nd_status nd_pack_scene(nd_ctx* c, const nd_uniforms& u, DevScene& s);
// A heightmap formula's image fields (h[0..8]) for image slot `slot`.
void nd_image_hoist(const nd_ctx* c, int slot, float* h);
// Some visible instance uses a formula beyond the four tuned ones.
bool nd_scene_ext(const DevScene& s);
nd_status nd_pack_job(nd_ctx* c, const nd_uniforms& u, const nd_march_opts& o, nd_rect rect, DevJob& j);
// Builds (or finds cached) the accelerator grids of the jobs that enable it and points their
// accel fields at them; launches on the context stream.
nd_status nd_accel_prepare(nd_ctx* c, DevJob* jobs, uint32_t n);
// Workgroups (one wave64 each) of a persistent pass whose kernel has residency class `res`,
// over about `units` claims of work (0 = unknown).
uint32_t nd_pass_grid(const nd_ctx* c, int res, uint64_t units);
