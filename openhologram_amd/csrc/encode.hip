// encode.hip -- 8-bit encodings of a fringe on gfx950.
//
//   oph_normalize_u8  ophNormalize (K/oph_common/oph_common.cpp:92-116):
//                     min/max, vertical flip, float scaling rounded by +0.5
//   oph_scale_u8      the scaling of HologramGenerator::Write_Result_image
//                     (DM/Hologram/src/HologramGenerator.cpp:528-565):
//                     double scaling truncated
// Both are two HBM passes: a min/max reduction (per-workgroup partials, then
// one workgroup; min and max are exact, so the result does not depend on the
// order) and the elementwise map.  The map reproduces the reference's
// operation order and rounding with explicitly rounded intrinsics, so the
// bytes equal the reference's.
#include <cfloat>

#include "oph_internal.h"

namespace {

constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 1024;

template <typename T>
struct Lim;
template <>
struct Lim<float> {
    __device__ static float lo() { return -FLT_MAX; }
    __device__ static float hi() { return FLT_MAX; }
};
template <>
struct Lim<double> {
    __device__ static double lo() { return -DBL_MAX; }
    __device__ static double hi() { return DBL_MAX; }
};

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void minmax_partial(const T* __restrict__ src, int64_t n,
                                                              T* __restrict__ part) {
    __shared__ T smin[RED_THREADS], smax[RED_THREADS];
    T mn = Lim<T>::hi(), mx = Lim<T>::lo();
    for (int64_t i = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; i < n; i += (int64_t)gridDim.x * RED_THREADS) {
        const T v = src[i];
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            smin[threadIdx.x] = smin[threadIdx.x + s] < smin[threadIdx.x] ? smin[threadIdx.x + s] : smin[threadIdx.x];
            smax[threadIdx.x] = smax[threadIdx.x + s] > smax[threadIdx.x] ? smax[threadIdx.x + s] : smax[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smin[0];
        part[2 * blockIdx.x + 1] = smax[0];
    }
}

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void minmax_final(T* part, int nblocks) {
    __shared__ T smin[RED_THREADS], smax[RED_THREADS];
    T mn = Lim<T>::hi(), mx = Lim<T>::lo();
    for (int i = threadIdx.x; i < nblocks; i += RED_THREADS) {
        mn = part[2 * i] < mn ? part[2 * i] : mn;
        mx = part[2 * i + 1] > mx ? part[2 * i + 1] : mx;
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int s = RED_THREADS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            smin[threadIdx.x] = smin[threadIdx.x + s] < smin[threadIdx.x] ? smin[threadIdx.x + s] : smin[threadIdx.x];
            smax[threadIdx.x] = smax[threadIdx.x + s] > smax[threadIdx.x] ? smax[threadIdx.x + s] : smax[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[0] = smin[0];
        part[1] = smax[0];
    }
}

// (unsigned char)(((v - mn) / (mx - mn)) * 255 + 0.5): float subtract,
// divide and multiply, then the double add of the literal 0.5, truncation.
__global__ void normalize_map(const float* __restrict__ src, unsigned char* __restrict__ dst, int nx, int ny,
                              const float* __restrict__ mm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nx * ny) return;
    const float mn = mm[0], mx = mm[1];
    const int x = (int)(i % nx), y = (int)(i / nx);
    unsigned char o = 0;
    if (mx > mn) {
        const float q = __fmul_rn(__fdiv_rn(__fsub_rn(src[i], mn), __fsub_rn(mx, mn)), 255.0f);
        o = (unsigned char)(int)__dadd_rn((double)q, 0.5);
    }
    dst[(int64_t)(ny - 1 - y) * nx + x] = o;
}

// (uint)((v - mn) / (mx - mn) * 255) in double
__global__ void scale_map(const double* __restrict__ src, unsigned char* __restrict__ dst, int64_t n,
                          const double* __restrict__ mm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double mn = mm[0], mx = mm[1];
    unsigned char o = 0;
    if (mx > mn) o = (unsigned char)(unsigned)__dmul_rn(__ddiv_rn(__dsub_rn(src[i], mn), __dsub_rn(mx, mn)), 255.0);
    dst[i] = o;
}

template <typename T>
int minmax(oph_ctx* ctx, const T* src, int64_t n, T* part) {
    const int nb = (int)std::min<int64_t>(RED_BLOCKS, (n + RED_THREADS - 1) / RED_THREADS);
    OPH_LAUNCH(ctx, "minmax_partial", minmax_partial<T>, dim3(nb), dim3(RED_THREADS), 0, src, n, part);
    OPH_LAUNCH(ctx, "minmax_final", minmax_final<T>, dim3(1), dim3(RED_THREADS), 0, part, nb);
    return OPH_OK;
}

// stages a host input (or uses the device pointer) and provides the output
template <typename T>
int encode_common(oph_ctx* ctx, const T* src, unsigned char* dst, int64_t n, int flags, const T** dsrc,
                  unsigned char** ddst, T** part) {
    const bool dev = flags & OPH_DEVICE_PTR;
    if ((flags & OPH_ASYNC) && !dev) return oph_fail(ctx, OPH_EINVAL, "OPH_ASYNC needs device pointers");
    *part = static_cast<T*>(oph_scratch(ctx, SLOT_AUX1, sizeof(T) * 2 * RED_BLOCKS));
    if (!*part) return OPH_ENOMEM;
    if (dev) {
        *dsrc = src;
        *ddst = dst;
        return OPH_OK;
    }
    char* buf = static_cast<char*>(oph_scratch(ctx, SLOT_IN, sizeof(T) * n + n + 256));
    if (!buf) return OPH_ENOMEM;
    OPH_HIP_CHECK(ctx, hipMemcpyAsync(buf, src, sizeof(T) * n, hipMemcpyHostToDevice, ctx->stream));
    *dsrc = reinterpret_cast<const T*>(buf);
    *ddst = reinterpret_cast<unsigned char*>(buf + ((sizeof(T) * n + 255) & ~(size_t)255));
    return OPH_OK;
}

}  // namespace

extern "C" int oph_normalize_u8(oph_ctx* ctx, const float* src, unsigned char* dst, int nx, int ny, int flags) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!src || !dst || nx <= 0 || ny <= 0) return oph_fail(ctx, OPH_EINVAL, "bad normalize arguments");
    const int64_t n = (int64_t)nx * ny;
    const float* s;
    unsigned char* d;
    float* part;
    if ((rc = encode_common(ctx, src, dst, n, flags, &s, &d, &part))) return rc;
    if ((rc = minmax(ctx, s, n, part))) return rc;
    OPH_LAUNCH(ctx, "normalize_map", normalize_map, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, nx, ny,
               part);
    if (!(flags & OPH_DEVICE_PTR))
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(dst, d, n, hipMemcpyDeviceToHost, ctx->stream));
    return oph_finish(ctx, flags);
}

extern "C" int oph_scale_u8(oph_ctx* ctx, const double* src, unsigned char* dst, int64_t n, int flags) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!src || !dst || n <= 0) return oph_fail(ctx, OPH_EINVAL, "bad scale arguments");
    const double* s;
    unsigned char* d;
    double* part;
    if ((rc = encode_common(ctx, src, dst, n, flags, &s, &d, &part))) return rc;
    if ((rc = minmax(ctx, s, n, part))) return rc;
    OPH_LAUNCH(ctx, "scale_map", scale_map, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d, n, part);
    if (!(flags & OPH_DEVICE_PTR))
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(dst, d, n, hipMemcpyDeviceToHost, ctx->stream));
    return oph_finish(ctx, flags);
}
