// oph_internal.h -- context, error plumbing and device helpers shared by the
// HIP translation units of liboph_hip.so.  gfx950 (CDNA4) only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "oph_hip.h"

// ---------------------------------------------------------------------------
// context

enum ScratchSlot {
    SLOT_PTS = 0,   // staged point arrays
    SLOT_OUT,       // staged output
    SLOT_IN,        // staged input
    SLOT_WORK0,     // FFT / pipeline work buffers
    SLOT_WORK1,
    SLOT_WORK2,
    SLOT_TABLE,     // small tables (twiddles, transfer phases)
    SLOT_TABLE2,
    SLOT_AUX0,
    SLOT_AUX1,
    SLOT_COUNT
};

struct LaunchRecord {
    const char* name;
    hipEvent_t a, b;
};

struct oph_ctx {
    int device = 0;
    // per-kernel timing with HIP events on this context's stream
    // (oph_profile_enable); events are pooled and reused
    int profiling = 0;
    std::vector<LaunchRecord> launches;
    std::vector<hipEvent_t> event_pool;
    size_t pool_used = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    void* slot_ptr[SLOT_COUNT] = {};
    size_t slot_size[SLOT_COUNT] = {};
    std::string err;
    int cu_count = 256;
    void* tw[24] = {};  // FFT twiddle tables by log2(length), radix-16 stages
    void* tw8[24] = {}; // the same for radix-8 stages
    void* twm[16] = {}; // mixed-radix tables (lengths 2^a 3^b 5^c) ...
    int twm_len[16] = {};  // ... and their lengths
    unsigned* wrp_flag = nullptr;          // device word set by an async WRP call whose lists overflowed
    std::vector<double> host_scratch;      // per-call host tables uploaded to the device
};

// Records an error message on the context and returns code.
int oph_fail(oph_ctx* ctx, int code, const char* fmt, ...);

// Grow-only device buffer owned by the context.  Returns nullptr (with the
// error recorded) when allocation fails.
void* oph_scratch(oph_ctx* ctx, int slot, size_t bytes);

#define OPH_HIP_CHECK(ctx, expr)                                                  \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess)                                                     \
            return oph_fail((ctx), OPH_EDEVICE, "%s: %s (%s:%d)", #expr,          \
                            hipGetErrorString(e_), __FILE__, __LINE__);           \
    } while (0)

#define OPH_CHECK_LAUNCH(ctx) OPH_HIP_CHECK(ctx, hipGetLastError())

// Event pair around one kernel launch when profiling is on.
hipEvent_t oph_prof_event(oph_ctx* ctx);
#define OPH_LAUNCH(ctx, name, kernel, grid, block, shmem, ...)                      \
    do {                                                                          \
        hipEvent_t pa_ = nullptr, pb_ = nullptr;                                  \
        if ((ctx)->profiling) {                                                   \
            pa_ = oph_prof_event(ctx);                                            \
            pb_ = oph_prof_event(ctx);                                            \
            if (pa_) hipEventRecord(pa_, (ctx)->stream);                          \
        }                                                                         \
        hipLaunchKernelGGL(kernel, grid, block, shmem, (ctx)->stream, __VA_ARGS__); \
        if (pb_) {                                                                \
            hipEventRecord(pb_, (ctx)->stream);                                   \
            (ctx)->launches.push_back(LaunchRecord{name, pa_, pb_});              \
        }                                                                         \
        OPH_CHECK_LAUNCH(ctx);                                                    \
    } while (0)

// Makes the context's device current.
int oph_enter(oph_ctx* ctx);
// Waits on the stream unless OPH_ASYNC; reports async errors.
int oph_finish(oph_ctx* ctx, int flags);

// Staging of one array argument: returns a device pointer holding `bytes`
// of `host` (copied when `host` is a host pointer) or `host` itself when the
// call passed device pointers.
const void* oph_stage_in(oph_ctx* ctx, int slot, const void* p, size_t bytes, int flags,
                         size_t offset = 0);

// ---------------------------------------------------------------------------
// device helpers

// sin/cos of 2*pi*t for t in turns, |t| <= 256: v_sin_f32 / v_cos_f32 take
// their argument in turns.
__device__ __forceinline__ float sin_turns(float t) { return __builtin_amdgcn_sinf(t); }
__device__ __forceinline__ float cos_turns(float t) { return __builtin_amdgcn_cosf(t); }

// Fractional part of a double, in [-0.5, 0.5].
__device__ __forceinline__ double frac_turns(double t) { return t - rint(t); }

// 1/sqrt(s) in double: hardware estimate refined by two Newton steps.
__device__ __forceinline__ double rsqrt_f64(double s) {
    double y = __builtin_amdgcn_rsq(s);
    double h = 0.5 * s;
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

struct f2 { float x, y; };

// Raw buffer resource over [p, p + bytes), bytes < 2^31 (gfx9 word 3).
// Buffer loads and stores are range-checked by the hardware: an offset at
// or beyond `bytes` reads zeros and writes nothing, so bounds are handled
// without branches, which keeps the compiler's memory wait counts exact
// across a loop (the counter is in order: a branch around a load or store
// makes the next wait conservative).
constexpr unsigned OPH_BUF_OOB = 0x80000000u;  // an offset beyond every buffer made here
__device__ __forceinline__ __amdgpu_buffer_rsrc_t oph_buf_rsrc(const void* p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}
// (the builtins' vectors are converted as a whole with __builtin_bit_cast;
// subscripting the returned vector lost its second element here)
typedef float oph_f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned oph_u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 oph_buf_load_f2(__amdgpu_buffer_rsrc_t r, unsigned off) {
    const oph_f32x2 v = __builtin_bit_cast(oph_f32x2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
    return make_float2(v.x, v.y);
}
// (soff: a wave-uniform byte offset, kept in a scalar register)
__device__ __forceinline__ float2 oph_buf_load_f2(__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
    const oph_f32x2 v = __builtin_bit_cast(oph_f32x2, __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0));
    return make_float2(v.x, v.y);
}
typedef float oph_f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned oph_u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 oph_buf_load_f4(__amdgpu_buffer_rsrc_t r, unsigned off) {
    const oph_f32x4 v = __builtin_bit_cast(oph_f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void oph_buf_store_f4(float4 x, __amdgpu_buffer_rsrc_t r, unsigned off) {
    const oph_f32x4 v = {x.x, x.y, x.z, x.w};
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(oph_u32x4, v), r, off, 0, 0);
}
__device__ __forceinline__ void oph_buf_store_f2(float2 x, __amdgpu_buffer_rsrc_t r, unsigned off) {
    const oph_f32x2 v = {x.x, x.y};
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(oph_u32x2, v), r, off, 0, 0);
}
__device__ __forceinline__ void oph_buf_store_f2(float2 x, __amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
    const oph_f32x2 v = {x.x, x.y};
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(oph_u32x2, v), r, off, soff, 0);
}

// Workgroup index remap for the 8 XCDs of MI355X: consecutive hardware ids go
// round-robin over XCDs; this returns a logical id such that each XCD receives
// a contiguous range of logical ids (neighbouring tiles then share that XCD's
// L2).  n = number of workgroups in the grid.
__device__ __forceinline__ int xcd_remap(int hw, int n) {
    const int nx = 8;
    const int per = n / nx;              // full ranges
    const int rem = n - per * nx;
    const int xcd = hw % nx, slot = hw / nx;
    if (slot >= per) return per * nx + xcd;  // tail (n not a multiple of 8)
    (void)rem;
    return xcd * per + slot;
}
