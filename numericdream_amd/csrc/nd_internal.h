// nd_internal.h - launch records shared between the C-ABI host code and the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include "nd_dev.h"

struct NdPassLaunch {
    const DevJob* jobs;          // device array of jobs
    const DevScene* shared;      // device scene shared by all jobs, or null
    const uint32_t* job_base;    // device: first ray slot of each job
    int n_jobs;
    uint32_t total_rays;
    DevCounters* ctr;
    DevHitBuffers hb;
    uint32_t grid_march;         // persistent workgroups (one wave64 each) of each pass
    uint32_t grid_surface;
    uint32_t grid_shadow;
    uint32_t chunk;
    uint32_t min_idle;
    uint32_t round_cap;
    uint32_t refill_period;
    int any_shadow;
    int surface;                 // run the normal / trap pass
    unsigned long long* job_done_cycles;  // per-job completion clock, or null
    uint32_t* ray_done;          // per ray slot completion tick (total_rays entries) when tracking
    int accel;                   // some job uses the empty-space accelerator
    int ext;                     // some instance uses a formula beyond the four tuned ones
    hipEvent_t ev[4];            // optional: before march / surface / shadow / after
};

hipError_t nd_launch_plain(const DevJob* d_job, int tiles_x, int tiles_y, DevCounters* ctr, bool ext, hipStream_t s);
hipError_t nd_launch_zero_counters(DevCounters* c, hipStream_t s);
hipError_t nd_launch_persistent(const NdPassLaunch& L, hipStream_t s);
// fs_simple colour and fs_aov albedo of every pixel of a batch after its geometry passes
// (scenes: one packed material scene per job, nd_mc_scene_bytes() each).
constexpr size_t ND_MC_SCENE_MAX_BYTES = 8192;  // upper bound of nd_mc_scene_bytes()
size_t nd_mc_scene_bytes();
void nd_mc_scene_pack(const nd_uniforms& u, void* out);
hipError_t nd_launch_gbuffer_color(const DevJob* jobs, const void* scenes, const uint32_t* job_base, int n_jobs,
                                   uint32_t total, bool ext, hipStream_t s);

// The radiance adjoint over the compacted hits of the batch just run (nd_mc.hip).
struct NdRadianceLaunch {
    const DevJob* jobs;
    const DevScene* shared;  // non-null when every job has the same scene
    const DevCounters* ctr;
    DevHitBuffers hb;
    const void* scenes;      // one packed material scene per job
    const float* target;     // 3 floats per pixel, or null (no loss)
    float* rgb;              // 3 floats per pixel, or null
    float* drgb;             // 18 floats per pixel ([c * 6 + k]), or null
    const uint64_t* pix_base;  // first pixel of each job
    double* acc;             // loss, 6 gradient components, pixels used
};
// ext: some job uses a formula beyond the four tuned ones (the kernel then reads each job's scene)
hipError_t nd_launch_radiance_adjoint(const NdRadianceLaunch& L, uint32_t grid, bool ext, hipStream_t s);
uint32_t nd_resident_radiance(bool ext);  // resident wave64 workgroups per CU of the radiance adjoint

// Resident wave64 workgroups per CU of each persistent kernel (blockDim 64, no dynamic LDS).
// "_MIXED" variants take the scene from each job, the others share one scene; "_EXT"
// variants also run the formulas beyond the four tuned ones (and take each job's scene).
enum NdResident {
    RES_MARCH, RES_MARCH_MIXED, RES_MARCH_ACC, RES_MARCH_ACC_MIXED, RES_MARCH_EXT, RES_MARCH_ACC_EXT,
    RES_SURFACE, RES_SURFACE_MIXED, RES_SURFACE_EXT,
    RES_SHADOW, RES_SHADOW_MIXED, RES_SHADOW_ACC, RES_SHADOW_ACC_MIXED, RES_SHADOW_EXT, RES_SHADOW_ACC_EXT,
    RES_ADJOINT, RES_ADJOINT_MIXED, RES_ADJOINT_EXT,
    RES_MC_EXTEND, RES_MC_SHADE, RES_MC_SHADOW, RES_MC_IND_MARCH, RES_MC_IND_SHADE,
    RES_MC_EXTEND_EXT, RES_MC_SHADE_EXT, RES_MC_SHADOW_EXT, RES_MC_IND_MARCH_EXT, RES_MC_IND_SHADE_EXT,
    RES_COUNT
};
void nd_resident_march(uint32_t* res);    // fills RES_MARCH .. RES_SHADOW_ACC_MIXED
void nd_resident_adjoint(uint32_t* res);  // RES_ADJOINT, RES_ADJOINT_MIXED
void nd_resident_mc(uint32_t* res);       // RES_MC_*
