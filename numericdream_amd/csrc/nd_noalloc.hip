// nd_noalloc.hip - nd_render_gbuffer_noalloc: the G-buffer pass without any device
// allocation, for callers that cannot allocate (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc).
//
// Every buffer beyond the caller's inputs and outputs - the job record, the shared scene, the
// pass counters and the persistent schedule's hit list - is carved from one file-scope device
// array of fixed size. Each GPU holds its own copy of it from module load until the process
// ends. Host code never passes the array's name to a kernel (on the host it is a stand-in, not
// a device address): the device address is looked up once per device with
// hipGetSymbolAddress, with that device current, and used everywhere instead.
#include <hip/hip_runtime.h>

#include <mutex>

#include "nd_ctx.h"

namespace {

constexpr size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes per hit slot of the persistent schedule (ray, job, t, opix, steps, shadow origin,
// direction and distance), as nd_batch_impl carves them, plus its rounding slack.
constexpr size_t kHitSlotBytes = 4 + 2 + 4 + 4 + 2 + 7 * 4;
constexpr size_t kHitSlots = (size_t)ND_NOALLOC_MAX_RAYS + (size_t)ND_NOALLOC_MAX_GRID * 64;
constexpr size_t kScratchBytes = round256(sizeof(DevJob)) + round256(sizeof(uint32_t)) + round256(sizeof(DevCounters)) +
                                 round256(sizeof(DevScene)) + round256(ND_MC_SCENE_MAX_BYTES) +
                                 kHitSlots * kHitSlotBytes + 16 * 256;

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// The scratch array. Never named by host code other than in the address lookup below.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

char* nd_noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        // the caller has made `device` current: the address is that device's copy
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        g_scratch_ptr[device] = (char*)p;
    }
    return g_scratch_ptr[device];
}

size_t nd_noalloc_scratch_bytes() { return kScratchBytes; }

extern "C" nd_status nd_render_gbuffer_noalloc(nd_ctx* c, const void* uniforms, size_t size, const nd_march_opts* opts,
                                               nd_rect rect, const nd_gbuffer_out* out, nd_batch_stats* stats) {
    if (!c) return ND_E_INVALID;
    if (!uniforms) return nd_fail(c, ND_E_INVALID, "null uniforms");
    if (size != sizeof(nd_uniforms)) return nd_fail(c, ND_E_INVALID, "uniforms_size %zu != %zu", size, sizeof(nd_uniforms));
    if (!out) return nd_fail(c, ND_E_INVALID, "null output set");
    if (!c->noalloc_scratch) return nd_fail(c, ND_E_DEVICE, "the scratch array of device %d has no device address", c->device);
    NdFixedWorkspace w{c->noalloc_scratch, kScratchBytes, ND_NOALLOC_MAX_RAYS, ND_NOALLOC_MAX_GRID};
    NdBatchExtra ex;
    ex.fixed = &w;
    nd_job j;
    j.uniforms = (const nd_uniforms*)uniforms;
    j.rect = rect;
    j.out = *out;
    return nd_batch_impl(c, &j, 1, opts, stats, &ex);
}
