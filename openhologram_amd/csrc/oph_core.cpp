// oph_core.cpp -- context lifetime, errors, device memory helpers, timing,
// and the KETI file readers of the C ABI.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "oph_internal.h"
#include "pipelines.h"

int oph_fail(oph_ctx* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

void* oph_scratch(oph_ctx* ctx, int slot, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (ctx->slot_size[slot] >= bytes) return ctx->slot_ptr[slot];
    if (ctx->slot_ptr[slot]) {
        hipStreamSynchronize(ctx->stream);
        hipFree(ctx->slot_ptr[slot]);
        ctx->slot_ptr[slot] = nullptr;
        ctx->slot_size[slot] = 0;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        hipGetLastError();
        oph_fail(ctx, OPH_ENOMEM, "device allocation of %zu bytes failed", bytes);
        return nullptr;
    }
    ctx->slot_ptr[slot] = p;
    ctx->slot_size[slot] = bytes;
    return p;
}

int oph_enter(oph_ctx* ctx) {
    if (!ctx) return OPH_EINVAL;
    OPH_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return OPH_OK;
}

int oph_finish(oph_ctx* ctx, int flags) {
    OPH_CHECK_LAUNCH(ctx);
    if (!(flags & OPH_ASYNC)) OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return OPH_OK;
}

const void* oph_stage_in(oph_ctx* ctx, int slot, const void* p, size_t bytes, int flags,
                         size_t offset) {
    if (!p) return nullptr;
    if (flags & OPH_DEVICE_PTR) return p;
    char* d = static_cast<char*>(oph_scratch(ctx, slot, offset + bytes));
    if (!d) return nullptr;
    if (hipMemcpyAsync(d + offset, p, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
        hipGetLastError();
        oph_fail(ctx, OPH_EDEVICE, "host-to-device copy of %zu bytes failed", bytes);
        return nullptr;
    }
    return d + offset;
}

hipEvent_t oph_prof_event(oph_ctx* ctx) {
    if (ctx->pool_used == ctx->event_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) {
            hipGetLastError();
            return nullptr;
        }
        ctx->event_pool.push_back(e);
    }
    return ctx->event_pool[ctx->pool_used++];
}

extern "C" {

int oph_profile_enable(oph_ctx* ctx, int on) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->launches.clear();
    ctx->pool_used = 0;
    ctx->profiling = on ? 1 : 0;
    return OPH_OK;
}

int oph_profile_get(oph_ctx* ctx, const char* kernel, int* count, double* total_ms) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!kernel) return oph_fail(ctx, OPH_EINVAL, "null kernel name");
    OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    int n = 0;
    double t = 0.0;
    for (const auto& r : ctx->launches) {
        if (std::strcmp(r.name, kernel) != 0) continue;
        float ms = 0.f;
        OPH_HIP_CHECK(ctx, hipEventElapsedTime(&ms, r.a, r.b));
        t += ms;
        ++n;
    }
    if (count) *count = n;
    if (total_ms) *total_ms = t;
    return OPH_OK;
}

const char* oph_version(void) { return "oph_hip 1.0 gfx950"; }

int oph_create(int device, oph_ctx** out) {
    if (!out) return OPH_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        hipGetLastError();
        return OPH_EDEVICE;
    }
    if (device < 0 || device >= n) return OPH_EINVAL;
    oph_ctx* c = new oph_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev_begin) != hipSuccess || hipEventCreate(&c->ev_end) != hipSuccess) {
        hipGetLastError();
        delete c;
        return OPH_EDEVICE;
    }
    // the context's asynchronous-WRP overflow word, cleared once here
    void* flag = nullptr;
    if (hipMalloc(&flag, sizeof(unsigned)) != hipSuccess || hipMemset(flag, 0, sizeof(unsigned)) != hipSuccess) {
        hipGetLastError();
        if (flag) hipFree(flag);
        hipEventDestroy(c->ev_begin);
        hipEventDestroy(c->ev_end);
        hipStreamDestroy(c->stream);
        delete c;
        return OPH_ENOMEM;
    }
    c->wrp_flag = static_cast<unsigned*>(flag);
    // the allocation-free WRP's static scratch: its address on this device
    if (oph_wrp_noalloc_setup(c, device) != OPH_OK) {
        hipFree(flag);
        hipEventDestroy(c->ev_begin);
        hipEventDestroy(c->ev_end);
        hipStreamDestroy(c->stream);
        delete c;
        return OPH_EDEVICE;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->cu_count = prop.multiProcessorCount;
    *out = c;
    return OPH_OK;
}

void oph_destroy(oph_ctx* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    if (ctx->stream) hipStreamSynchronize(ctx->stream);
    for (int i = 0; i < SLOT_COUNT; ++i)
        if (ctx->slot_ptr[i]) hipFree(ctx->slot_ptr[i]);
    for (void* t : ctx->tw)
        if (t) hipFree(t);
    for (void* t : ctx->tw8)
        if (t) hipFree(t);
    for (void* t : ctx->twm)
        if (t) hipFree(t);
    if (ctx->wrp_flag) hipFree(ctx->wrp_flag);
    for (hipEvent_t e : ctx->event_pool) hipEventDestroy(e);
    if (ctx->ev_begin) hipEventDestroy(ctx->ev_begin);
    if (ctx->ev_end) hipEventDestroy(ctx->ev_end);
    if (ctx->stream) hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* oph_last_error(const oph_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int oph_sync(oph_ctx* ctx) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->wrp_flag) {
        unsigned f = 0;
        OPH_HIP_CHECK(ctx, hipMemcpy(&f, ctx->wrp_flag, sizeof f, hipMemcpyDeviceToHost));
        if (f) {
            const unsigned zero = 0;
            OPH_HIP_CHECK(ctx, hipMemcpy(ctx->wrp_flag, &zero, sizeof zero, hipMemcpyHostToDevice));
            return oph_fail(ctx, OPH_ENOMEM, "an asynchronous WRP call overflowed its tile lists and left its "
                                             "output unwritten; repeat it synchronously");
        }
    }
    return OPH_OK;
}

int oph_device_alloc(oph_ctx* ctx, size_t bytes, void** dptr) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!dptr) return oph_fail(ctx, OPH_EINVAL, "null output pointer");
    if (hipMalloc(dptr, bytes ? bytes : 1) != hipSuccess) {
        hipGetLastError();
        return oph_fail(ctx, OPH_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
    }
    return OPH_OK;
}

int oph_device_free(oph_ctx* ctx, void* dptr) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    OPH_HIP_CHECK(ctx, hipFree(dptr));
    return OPH_OK;
}

int oph_copy_to_device(oph_ctx* ctx, void* dst, const void* src, size_t bytes) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    OPH_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return OPH_OK;
}

int oph_copy_to_host(oph_ctx* ctx, void* dst, const void* src, size_t bytes) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    OPH_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return OPH_OK;
}

// Plain streaming copy, 16 B per lane per iteration (grid-stride).  Besides
// serving oph_device_copy it is the calibration kernel for hardware byte
// counters: it reads and writes exactly `n16` * 16 bytes.
__global__ void __launch_bounds__(256) copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                   size_t n16) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

int oph_device_copy(oph_ctx* ctx, void* dst, const void* src, size_t bytes) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    const bool aligned = (((uintptr_t)dst | (uintptr_t)src) & 15) == 0;
    const size_t n16 = aligned ? bytes / 16 : 0;
    if (n16) {
        const unsigned grid = (unsigned)std::min<size_t>((n16 + 255) / 256, (size_t)std::max(ctx->cu_count, 1) * 16);
        OPH_LAUNCH(ctx, "copy_kernel", copy_kernel, dim3(grid), dim3(256), 0, (const uint4*)src, (uint4*)dst, n16);
    }
    if (bytes > n16 * 16)
        OPH_HIP_CHECK(ctx, hipMemcpyAsync((char*)dst + n16 * 16, (const char*)src + n16 * 16, bytes - n16 * 16,
                                          hipMemcpyDeviceToDevice, ctx->stream));
    OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return OPH_OK;
}

int oph_device_zero(oph_ctx* ctx, void* dptr, size_t bytes) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    OPH_HIP_CHECK(ctx, hipMemsetAsync(dptr, 0, bytes, ctx->stream));
    OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return OPH_OK;
}

int oph_timer_begin(oph_ctx* ctx) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    OPH_HIP_CHECK(ctx, hipEventRecord(ctx->ev_begin, ctx->stream));
    return OPH_OK;
}

int oph_timer_end(oph_ctx* ctx, float* ms) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    OPH_HIP_CHECK(ctx, hipEventRecord(ctx->ev_end, ctx->stream));
    OPH_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev_end));
    float t = 0.f;
    OPH_HIP_CHECK(ctx, hipEventElapsedTime(&t, ctx->ev_begin, ctx->ev_end));
    if (ms) *ms = t;
    return OPH_OK;
}

// KETI spec: 17 lines `title = value`; only the line count and the value
// field matter (ophgen_pointcloud.cpp:125-154).  Values parsed with strtod so
// that the wavelength and pitches keep their decimal value in double.
int oph_spec_from_keti_config(const char* path, oph_spec* spec) {
    if (!path || !spec) return OPH_EINVAL;
    FILE* f = std::fopen(path, "r");
    if (!f) return OPH_EIO;
    std::vector<std::string> vals;
    char line[512];
    while (std::fgets(line, sizeof line, f)) {
        char title[256] = {0}, value[256] = {0};
        if (std::sscanf(line, "%255s = %255s", title, value) < 2) value[0] = 0;
        vals.emplace_back(value);
    }
    std::fclose(f);
    if (vals.size() != 17) return OPH_EIO;
    auto d = [&](int i) { return std::strtod(vals[i].c_str(), nullptr); };
    spec->scale[0] = d(0);
    spec->scale[1] = d(1);
    spec->scale[2] = d(2);
    spec->offset_z = d(3);
    spec->pitch_x = d(4);
    spec->pitch_y = d(5);
    spec->nx = std::atoi(vals[6].c_str());
    spec->ny = std::atoi(vals[7].c_str());
    spec->lambda = d(14);
    spec->tilt_deg[0] = d(15);
    spec->tilt_deg[1] = d(16);
    return OPH_OK;
}

int oph_keti_config_fields(const char* path, char values[17][64]) {
    if (!path || !values) return OPH_EINVAL;
    FILE* f = std::fopen(path, "r");
    if (!f) return OPH_EIO;
    int i = 0;
    char line[512];
    while (std::fgets(line, sizeof line, f)) {
        char title[256] = {0}, value[256] = {0};
        if (std::sscanf(line, "%255s = %255s", title, value) < 2) value[0] = 0;
        if (i < 17) {
            std::strncpy(values[i], value, 63);
            values[i][63] = 0;
        }
        ++i;
    }
    std::fclose(f);
    return i == 17 ? OPH_OK : OPH_EIO;
}

// KETI .dat point cloud, ophgen_pointcloud.cpp:13-47: first token P, then per
// point `idx x y z phase amplitude`; an index that differs from the line
// number fails the load (:32-42).
int oph_load_keti_points(const char* path, double* xyzpa, int64_t cap, int64_t* n) {
    if (!path || !n) return OPH_EINVAL;
    FILE* f = std::fopen(path, "r");
    if (!f) return OPH_EIO;
    long long np = 0;
    if (std::fscanf(f, "%lld", &np) != 1 || np < 0) {
        std::fclose(f);
        return OPH_EIO;
    }
    for (long long i = 0; i < np; ++i) {
        long long idx;
        double v[5];
        if (std::fscanf(f, "%lld %lf %lf %lf %lf %lf", &idx, &v[0], &v[1], &v[2], &v[3], &v[4]) != 6 ||
            idx != i) {
            std::fclose(f);
            return OPH_EIO;
        }
        if (xyzpa && i < cap)
            for (int c = 0; c < 5; ++c) xyzpa[5 * i + c] = v[c];
    }
    std::fclose(f);
    *n = np;
    return OPH_OK;
}

}  // extern "C"
