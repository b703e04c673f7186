// encode_field.hip -- the MATLAB toolbox's hologram encodings of a complex
// field on gfx950 (OpenHolo_Encoding/fn_encoding.m:1-98).
//
//   SSB         single sideband: centred spectrum, one half zeroed
//               (:29-47), inverse, real part scaled to [0, 1]
//   NI          numerical interference with a plane reference of amplitude
//               max|U| (:49-54), scaled by its maximum
//   P2A         angle(U) + pi/(2 pi) as the reference writes it (:56-58):
//               angle + 0.5
//   REAL        Re(U) (:20-21)
//   BURCKHARDT  three non-negative components per pixel (:60-83)
//   TWO_PHASE   two phases angle + pi +- acos(|U|/max|U|) per pixel (:85-98)
// The reference's quirks are kept: the sideband block starts at 1-based
// index floor(N/2) (so 'left' also drops the DC column and the one below
// it), P2A's operator precedence, and Burckhardt's second component of each
// branch without the amplitude factor.
//
// fftshift(fft2(ifftshift(.))) and its inverse on even sizes: the shifts
// are circular shifts by N/2, which the forward and inverse transforms
// undo between themselves, so SSB is IFFT(M FFT(U)) with the mask M in FFT
// order: 'left' zeroes frequencies -1 .. N/2-1 (FFT indices 0 .. N/2-1 and
// N-1), 'right' zeroes -N/2 .. -1 (indices N/2 .. N-1); 'top'/'bottom' the
// same on rows.
//
// Reductions (max |U|^2, min/max of a float array) are exact in any order;
// every pass is one HBM sweep.
#include <cfloat>

#include "fft_kernels.h"
#include "oph_internal.h"

namespace {

constexpr int RT = 256;      // threads per reduction workgroup
constexpr int RB = 1024;     // reduction workgroups (at most)

// part[2b] = min, part[2b+1] = max of f(i) over this workgroup's stride
template <typename F>
__device__ void reduce_minmax(F f, int64_t n, float* part) {
    __shared__ float smin[RT], smax[RT];
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (int64_t i = (int64_t)blockIdx.x * RT + threadIdx.x; i < n; i += (int64_t)gridDim.x * RT) {
        const float v = f(i);
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int s = RT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smin[0];
        part[2 * blockIdx.x + 1] = smax[0];
    }
}

__global__ __launch_bounds__(RT) void abs2_minmax_partial(const float2* __restrict__ u, int64_t n, float* part) {
    reduce_minmax([=](int64_t i) { const float2 v = u[i]; return fmaf(v.x, v.x, v.y * v.y); }, n, part);
}

__global__ __launch_bounds__(RT) void minmax_partial_f(const float* __restrict__ s, int64_t n, float* part) {
    reduce_minmax([=](int64_t i) { return s[i]; }, n, part);
}

// part[0] = min, part[1] = max over the nb partials (pairs)
__global__ __launch_bounds__(RT) void minmax_combine(float* part, int nb) {
    __shared__ float smin[RT], smax[RT];
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (int i = threadIdx.x; i < nb; i += RT) {
        mn = fminf(mn, part[2 * i]);
        mx = fmaxf(mx, part[2 * i + 1]);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int s = RT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[0] = smin[0];
        part[1] = smax[0];
    }
}

// zero the stop band of a spectrum in FFT order
__global__ void ssb_mask(float2* a, int nx, int ny, int passband) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nx * ny) return;
    const int c = (int)(i % nx), r = (int)(i / nx);
    bool stop;
    switch (passband) {
        case OPH_PASS_LEFT: stop = c < nx / 2 || c == nx - 1; break;
        case OPH_PASS_RIGHT: stop = c >= nx / 2; break;
        case OPH_PASS_TOP: stop = r < ny / 2 || r == ny - 1; break;
        default: stop = r >= ny / 2; break;
    }
    if (stop) a[i] = make_float2(0.f, 0.f);
}

__global__ void real_part(const float2* __restrict__ a, float* __restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i].x;
}

// out = (v - min)/(max - min)  (SSB: realPositive / max(realPositive))
__global__ void shift_scale(float* out, int64_t n, const float* __restrict__ mm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (out[i] - mm[0]) / (mm[1] - mm[0]);
}

// NI first pass: |U + max|U||^2 (mm2[1] = max |U|^2)
__global__ void ni_map(const float2* __restrict__ u, float* __restrict__ out, int64_t n, const float* __restrict__ mm2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float ref = sqrtf(mm2[1]);
    const float2 v = u[i];
    const float re = v.x + ref;
    out[i] = fmaf(re, re, v.y * v.y);
}

// out /= max(out)
__global__ void div_max(float* out, int64_t n, const float* __restrict__ mm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = out[i] / mm[1];
}

__global__ void p2a_real_map(const float2* __restrict__ u, float* __restrict__ out, int64_t n, int p2a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 v = u[i];
    out[i] = p2a ? atan2f(v.y, v.x) + 0.5f : v.x;
}

// Burckhardt (fn_encoding.m:60-83) and two-phase (:85-98), one pixel per
// thread; mm2[1] = max |U|^2.  angle(U/m) = angle(U) for m > 0.
__global__ void burckhardt_map(const float2* __restrict__ u, float* __restrict__ out, int nx, int64_t n,
                               const float* __restrict__ mm2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float m = sqrtf(mm2[1]);
    const float2 v = u[i];
    const float re = v.x / m, im = v.y / m;
    const float am = sqrtf(fmaf(re, re, im * im));
    const float ph = atan2f(im, re) + 3.14159265358979323846f;
    constexpr float third = 2.09439510239319549f, two_third = 4.18879020478639098f, two_pi = 6.28318530717958648f;
    constexpr float inv_s3 = 0.57735026918962576f;
    float a1 = 0.f, a2 = 0.f, a3 = 0.f;
    if (ph >= 0.f && ph < third) {
        a1 = am * (cosf(ph) + sinf(ph) * inv_s3);
        a2 = 2.f * sinf(ph) * inv_s3;
    } else if (ph >= third && ph < two_third) {
        const float q = ph - third;
        a2 = am * (cosf(q) + sinf(q) * inv_s3);
        a3 = 2.f * sinf(q) * inv_s3;
    } else if (ph >= two_third && ph < two_pi) {
        const float q = ph - two_third;
        a3 = am * (cosf(q) + sinf(q) * inv_s3);
        a1 = 2.f * sinf(q) * inv_s3;
    }
    const int64_t r = i / nx, c = i % nx;
    float* o = out + r * 3 * nx + 3 * c;
    o[0] = a1;
    o[1] = a2;
    o[2] = a3;
}

__global__ void two_phase_map(const float2* __restrict__ u, float* __restrict__ out, int nx, int64_t n,
                              const float* __restrict__ mm2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float m = sqrtf(mm2[1]);
    const float2 v = u[i];
    const float re = v.x / m, im = v.y / m;
    const float am = fminf(sqrtf(fmaf(re, re, im * im)), 1.f);
    const float ph = atan2f(im, re) + 3.14159265358979323846f;
    const float d = acosf(am);
    const int64_t r = i / nx, c = i % nx;
    float* o = out + r * 2 * nx + 2 * c;
    o[0] = ph + d;
    o[1] = ph - d;
}

int reduce(oph_ctx* ctx, const char* name, int which, const void* src, int64_t n, float* part) {
    const int nb = (int)std::min<int64_t>(RB, (n + RT - 1) / RT);
    if (which == 0)
        OPH_LAUNCH(ctx, name, abs2_minmax_partial, dim3(nb), dim3(RT), 0, static_cast<const float2*>(src), n, part);
    else
        OPH_LAUNCH(ctx, name, minmax_partial_f, dim3(nb), dim3(RT), 0, static_cast<const float*>(src), n, part);
    OPH_LAUNCH(ctx, "minmax_combine", minmax_combine, dim3(1), dim3(RT), 0, part, nb);
    return OPH_OK;
}

}  // namespace

extern "C" int oph_encode_field(oph_ctx* ctx, const oph_c64* field, int nx, int ny, int method, int passband,
                                float* out, int flags) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!field || !out || nx <= 0 || ny <= 0) return oph_fail(ctx, OPH_EINVAL, "bad encode arguments");
    if (method < OPH_ENC_SSB || method > OPH_ENC_TWO_PHASE) return oph_fail(ctx, OPH_EINVAL, "unknown method %d", method);
    if (method == OPH_ENC_SSB) {
        if (passband < OPH_PASS_LEFT || passband > OPH_PASS_BOTTOM)
            return oph_fail(ctx, OPH_EINVAL, "unknown passband %d", passband);
        if (!fftk::supported(nx) || !fftk::supported(ny))
            return oph_fail(ctx, OPH_ENOTSUP, "SSB size %dx%d: each side must be a power of two in 16..8192", nx, ny);
    }
    if ((flags & OPH_ASYNC) && !(flags & OPH_DEVICE_PTR))
        return oph_fail(ctx, OPH_EINVAL, "OPH_ASYNC needs device pointers");
    const int64_t n = (int64_t)nx * ny;
    const int64_t nout = n * (method == OPH_ENC_BURCKHARDT ? 3 : method == OPH_ENC_TWO_PHASE ? 2 : 1);
    const float2* u = reinterpret_cast<const float2*>(field);
    float* o = out;
    if (!(flags & OPH_DEVICE_PTR)) {
        float2* du = static_cast<float2*>(oph_scratch(ctx, SLOT_IN, n * sizeof(float2)));
        o = static_cast<float*>(oph_scratch(ctx, SLOT_OUT, nout * sizeof(float)));
        if (!du || !o) return OPH_ENOMEM;
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(du, field, n * sizeof(float2), hipMemcpyHostToDevice, ctx->stream));
        u = du;
    }
    // two reductions' partials (NI reads the first's result during the second)
    float* part = static_cast<float*>(oph_scratch(ctx, SLOT_AUX1, sizeof(float) * 4 * RB));
    if (!part) return OPH_ENOMEM;
    float* part2 = part + 2 * RB;
    const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
    switch (method) {
        case OPH_ENC_SSB: {
            float2* w = static_cast<float2*>(oph_scratch(ctx, SLOT_WORK0, n * sizeof(float2)));
            if (!w) return OPH_ENOMEM;
            OPH_HIP_CHECK(ctx, hipMemcpyAsync(w, u, n * sizeof(float2), hipMemcpyDeviceToDevice, ctx->stream));
            if ((rc = oph_fft_rows(ctx, nx, -1, w, ny, nx, 1.f))) return rc;
            if ((rc = oph_fft_cols(ctx, ny, -1, w, nx, nx, 1.f))) return rc;
            OPH_LAUNCH(ctx, "ssb_mask", ssb_mask, grid, blk, 0, w, nx, ny, passband);
            if ((rc = oph_fft_rows(ctx, nx, +1, w, ny, nx, 1.f))) return rc;
            if ((rc = oph_fft_cols(ctx, ny, +1, w, nx, nx, (float)(1.0 / (double)n)))) return rc;
            OPH_LAUNCH(ctx, "real_part", real_part, grid, blk, 0, w, o, n);
            if ((rc = reduce(ctx, "minmax_partial_f", 1, o, n, part))) return rc;
            OPH_LAUNCH(ctx, "shift_scale", shift_scale, grid, blk, 0, o, n, part);
            break;
        }
        case OPH_ENC_NI:
            if ((rc = reduce(ctx, "abs2_minmax_partial", 0, u, n, part))) return rc;
            OPH_LAUNCH(ctx, "ni_map", ni_map, grid, blk, 0, u, o, n, part);
            if ((rc = reduce(ctx, "minmax_partial_f", 1, o, n, part2))) return rc;
            OPH_LAUNCH(ctx, "div_max", div_max, grid, blk, 0, o, n, part2);
            break;
        case OPH_ENC_P2A:
        case OPH_ENC_REAL:
            OPH_LAUNCH(ctx, "p2a_real_map", p2a_real_map, grid, blk, 0, u, o, n, method == OPH_ENC_P2A ? 1 : 0);
            break;
        case OPH_ENC_BURCKHARDT:
            if ((rc = reduce(ctx, "abs2_minmax_partial", 0, u, n, part))) return rc;
            OPH_LAUNCH(ctx, "burckhardt_map", burckhardt_map, grid, blk, 0, u, o, nx, n, part);
            break;
        default:
            if ((rc = reduce(ctx, "abs2_minmax_partial", 0, u, n, part))) return rc;
            OPH_LAUNCH(ctx, "two_phase_map", two_phase_map, grid, blk, 0, u, o, nx, n, part);
            break;
    }
    if (!(flags & OPH_DEVICE_PTR))
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(out, o, nout * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    return oph_finish(ctx, flags);
}
