// Test harness: a CPU emulation of the HIP subset the library's sources use, so that its
// kernels (the same source files, built with g++) run on the host. One OS thread; every lane
// is a coroutine; wave collectives (ballot, readfirstlane, shuffles) and __syncthreads are
// resolved by the scheduler in emu.cpp; device memory is host memory with a guard page after
// each allocation; atomics are plain read-modify-writes (one thread). Host-side math stands
// in for the hardware's approximate rcp/sqrt/exp2/log2, so results match the GPU's to FP32
// rounding, not bit for bit. Used by tests/test_simt_emulation.py only.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <type_traits>
#include <vector>

#define __host__
#define __device__
#define __global__
#define __shared__ static
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) ((const void*)&(x))
#define __HIP_MEMORY_SCOPE_AGENT 0

struct dim3 {
    uint32_t x, y, z;
    constexpr dim3(uint32_t a = 1, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};
struct EmuLane;
struct EmuIdx { dim3 tid, bid; int lane; };
extern EmuIdx* emu_cur;
extern dim3 emu_grid_dim, emu_block_dim;
#define threadIdx (emu_cur->tid)
#define blockIdx (emu_cur->bid)
#define blockDim emu_block_dim
#define gridDim emu_grid_dim

enum { EMU_BALLOT = 0, EMU_FIRST = 1, EMU_SHFL = 2, EMU_SYNC = 3, EMU_SHFL_UP = 4, EMU_SHFL_IDX = 5 };
uint64_t emu_collective(int kind, uint64_t v, int arg, const char* file, int line);
uint64_t emu_clock();

#define __builtin_amdgcn_ballot_w64(p) emu_collective(EMU_BALLOT, (p) ? 1u : 0u, 0, __FILE__, __LINE__)
#define __builtin_amdgcn_readfirstlane(v) ((int)(uint32_t)emu_collective(EMU_FIRST, (uint32_t)(v), 0, __FILE__, __LINE__))
#define __syncthreads() (void)emu_collective(EMU_SYNC, 0, 0, __FILE__, __LINE__)
template <class T>
inline T emu_shfl(int kind, T v, int off, const char* f, int l) {
    static_assert(sizeof(T) <= 8, "shfl");
    uint64_t b = 0;
    std::memcpy(&b, &v, sizeof(T));
    b = emu_collective(kind, b, off, f, l);
    T r;
    std::memcpy(&r, &b, sizeof(T));
    return r;
}
#define __shfl_down(v, off, ...) emu_shfl(EMU_SHFL, (v), (off), __FILE__, __LINE__)
#define __shfl_up(v, off, ...) emu_shfl(EMU_SHFL_UP, (v), (off), __FILE__, __LINE__)
#define __shfl(v, src, ...) emu_shfl(EMU_SHFL_IDX, (v), (src), __FILE__, __LINE__)
inline float __frsqrt_rn(float x) { return 1.0f / std::sqrt(x); }
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
inline long long min(long long a, long long b) { return a < b ? a : b; }
inline long long max(long long a, long long b) { return a > b ? a : b; }
inline unsigned long long min(unsigned long long a, unsigned long long b) { return a < b ? a : b; }
inline unsigned long long max(unsigned long long a, unsigned long long b) { return a > b ? a : b; }
inline float min(float a, float b) { return std::fmin(a, b); }
inline float max(float a, float b) { return std::fmax(a, b); }
inline int __lane_id() { return emu_cur->lane; }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline uint64_t wall_clock64() { return emu_clock(); }

template <class T> inline T atomicAdd(T* p, T v) { T o = *p; *p = o + v; return o; }
template <class T> inline T atomicMax(T* p, T v) { T o = *p; if (v > o) *p = v; return o; }
template <class T> inline T __hip_atomic_load(const T* p, int, int) { return *p; }

inline float emu_med3(float a, float b, float c) { return std::max(std::min(a, b), std::min(std::max(a, b), c)); }
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
#define __builtin_amdgcn_sqrtf(x) std::sqrt((float)(x))
#define __builtin_amdgcn_fmed3f(a, b, c) emu_med3((a), (b), (c))
#define __builtin_amdgcn_exp2f(x) std::exp2((float)(x))
#define __builtin_amdgcn_logf(x) std::log2((float)(x))
inline uint32_t __float_as_uint(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float __uint_as_float(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline float __fmul_rn(float a, float b) { return a * b; }
inline float __fadd_rn(float a, float b) { return a + b; }

// ---- runtime
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorMemoryAllocation = 2,
       hipErrorNotSupported = 801 };
typedef void* hipStream_t;
typedef void* hipEvent_t;
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
typedef void* hipGraphNode_t;
enum hipMemcpyKind { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
enum hipStreamCaptureStatus { hipStreamCaptureStatusNone = 0, hipStreamCaptureStatusActive = 1 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
enum { hipStreamNonBlocking = 1, hipDeviceAttributeWallClockRate = 1 };
struct hipDeviceProp_t { int multiProcessorCount; size_t totalGlobalMem; char gcnArchName[256]; char name[256]; };

hipError_t hipMalloc(void** p, size_t n);
template <class T> inline hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
hipError_t hipFree(void* p);
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { if (n) std::memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t = nullptr) { if (n) std::memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t = nullptr) { std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemset(void* d, int v, size_t n) { std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "emulated"; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int);
inline hipError_t hipDeviceGetAttribute(int* v, int, int) { *v = 100000; return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (void*)0x1; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t = nullptr);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus* s) { *s = hipStreamCaptureStatusNone; return hipSuccess; }
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) { return hipErrorNotSupported; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) { *g = nullptr; return hipErrorNotSupported; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t, hipGraphNode_t*, char*, size_t) { *e = nullptr; return hipErrorNotSupported; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
hipError_t hipGetSymbolAddress(void** p, const void* sym);
template <class K> inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, K, int, size_t) { *n = 8; return hipSuccess; }

void emu_run(dim3 grid, dim3 block, const char* name, std::function<void()> body);
void emu_check_arg(const void*);
template <class T> inline void emu_check(const T&) {}
struct DevScene;
struct DevJob;
void emu_check_scene(const DevScene* s);
void emu_check_job(const DevJob* j);
inline void emu_check(const DevScene* s) { emu_check_scene(s); }
inline void emu_check(DevScene* s) { emu_check_scene(s); }
inline void emu_check(const DevJob* j) { emu_check_job(j); }
inline void emu_check(DevJob* j) { emu_check_job(j); }
template <class K, class... A>
inline void emu_launch(const char* name, K kernel, dim3 grid, dim3 block, A... args) {
    (emu_check(args), ...);
    emu_run(grid, block, name, [=]() { kernel(args...); });
}
#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) emu_launch(#kernel, kernel, dim3(grid), dim3(block), __VA_ARGS__)
