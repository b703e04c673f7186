// recon.hip -- numerical reconstruction of a depth-map fringe through an eye
// model on gfx950.
//
// Replaces HologramGenerator::ReconstructImage,
// Test_Propagation_to_Eye_Pupil and Reconstruction
// (DM/Hologram/src/HologramGenerator_CPU.cpp:415-614).  With F = (nx px,
// ny py) the hologram size, the eye-pupil plane has pitch
// pe = lambda f / F and the pupil-plane coordinates of pixel (x, y) are
//   xe = -nx pe_x/2 + pe_x x,   ye = (ny pe_y/2 - pe_y) - pe_y y   (CPU.cpp:521-522)
//   hh   = fftshift(FFT2(fftshift(fringe)))                        (:512)
//   hh_c = hh exp(i pi (xe^2 + ye^2)/(lambda f))                   (:524-530)
//   hh_s = circshift(hh_c, -round(cx/pe_x), round(cy/pe_y))         (:560-563)
//   f_eye = L (f - d)/(L + f - d),  f_eff = f_eye L/(f_eye - L)     (:565-566)
//   hh_e = hh_s exp(i pi (xe^2 + ye^2)/(lambda f_eff))
//          [r < |lambda f_eff/(2 max(pe))|] [r < pupil/2]           (:570-584)
//   out  = |fftshift(FFT2(fftshift(hh_e)))|                          (:586-603;
//          the retina chirp has unit modulus and drops out of |.|)
// For even sizes fftshift(FFT2(fftshift(a)))[k] = (-1)^(kx+ky+nx/2+ny/2)
// FFT2((-1)^(x+y) a)[k], so each shifted transform is a sign modulation
// fused into the neighbouring elementwise kernel plus a plain FFT (the sign
// after the second transform also has unit modulus).  Chirp phases (up to
// ~1e4 rad) are reduced to turns in FP64, the FFTs run in FP32.
#include <cmath>

#include "fft_kernels.h"
#include "oph_internal.h"

namespace {

struct EyeArgs {
    int nx, ny;
    double pex, pey;          // pupil-plane pitch
    double ex0, ey0;          // xe of x = 0, ye of y = 0
    double c_field, c_eye;    // 1/(2 lambda f), 1/(2 lambda f_eff): turns per m^2
    double r_alias, r_pupil;  // mask radii, compared as sqrt(r^2) < radius (CPU.cpp:580-581)
    int shx, shy;             // circshift amounts (reference sign convention)
};

__global__ void recon_pre(const double* __restrict__ fringe, float2* __restrict__ a, int nx, int ny) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)nx * ny) return;
    const int x = (int)(i % nx), y = (int)(i / nx);
    const float v = (float)fringe[i];
    a[i] = make_float2(((x + y) & 1) ? -v : v, 0.f);
}

__device__ __forceinline__ float2 chirp(double xe, double ye, double c) {
    const float t = (float)frac_turns((xe * xe + ye * ye) * c);
    return make_float2(cos_turns(t), sin_turns(t));
}

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// hh_e at (x, y) from the first transform A, with the sign modulation of the
// second shifted transform applied.
__global__ void recon_eye(const float2* __restrict__ A, float2* __restrict__ b, EyeArgs e) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)e.nx * e.ny) return;
    const int x = (int)(p % e.nx), y = (int)(p / e.nx);
    // circshift: out[(i + shx) mod nx][(j + shy) mod ny] = in[i][j]
    int i = (x - e.shx) % e.nx, j = (y - e.shy) % e.ny;
    if (i < 0) i += e.nx;
    if (j < 0) j += e.ny;
    const double xe = e.ex0 + e.pex * x, ye = e.ey0 - e.pey * y;
    const double r = sqrt(xe * xe + ye * ye);
    float2 v = make_float2(0.f, 0.f);
    if (r < e.r_alias && r < e.r_pupil) {
        float2 h = A[(int64_t)j * e.nx + i];
        if ((i + j + e.nx / 2 + e.ny / 2) & 1) h = make_float2(-h.x, -h.y);
        const double xi = e.ex0 + e.pex * i, yj = e.ey0 - e.pey * j;
        h = cmulf(h, chirp(xi, yj, e.c_field));
        v = cmulf(h, chirp(xe, ye, e.c_eye));
        if ((x + y) & 1) v = make_float2(-v.x, -v.y);
    }
    b[p] = v;
}

__global__ void recon_mag(const float2* __restrict__ b, double* __restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 v = b[i];
    out[i] = sqrt((double)v.x * v.x + (double)v.y * v.y);
}

}  // namespace

extern "C" int oph_dm_reconstruct(oph_ctx* ctx, const oph_eye_spec* s, const double* fringe, int nx, int ny,
                                  double* out, int flags) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!s || !fringe || !out) return oph_fail(ctx, OPH_EINVAL, "null argument");
    if (!fftk::supported(nx) || !fftk::supported(ny))
        return oph_fail(ctx, OPH_ENOTSUP, "reconstruction size %dx%d: powers of two in 16..8192", nx, ny);
    if (!(s->lambda > 0) || !(s->pitch_x > 0) || !(s->pitch_y > 0) || s->field_lens == 0 || s->eye_length == 0)
        return oph_fail(ctx, OPH_EINVAL, "bad eye spec");
    if ((flags & OPH_ASYNC) && !(flags & OPH_DEVICE_PTR))
        return oph_fail(ctx, OPH_EINVAL, "OPH_ASYNC needs device pointers");
    const int64_t n = (int64_t)nx * ny;
    const double f = s->field_lens, L = s->eye_length, lam = s->lambda;
    EyeArgs e;
    e.nx = nx;
    e.ny = ny;
    e.pex = lam * f / (nx * s->pitch_x);
    e.pey = lam * f / (ny * s->pitch_y);
    e.ex0 = -(e.pex * nx) / 2.0;
    e.ey0 = (e.pey * ny) / 2.0 - e.pey;
    const double f_eye = L * (f - s->focus_distance) / (L + (f - s->focus_distance));
    const double f_eff = f_eye * L / (f_eye - L);
    e.c_field = 1.0 / (2.0 * lam * f);
    e.c_eye = 1.0 / (2.0 * lam * f_eff);
    e.r_alias = fabs(lam * f_eff / (2.0 * fmax(e.pex, e.pey)));
    e.r_pupil = s->eye_pupil_diameter / 2.0;
    e.shx = -(int)lround(s->eye_center_x / e.pex);
    e.shy = (int)lround(s->eye_center_y / e.pey);
    if (!std::isfinite(f_eff) || !std::isfinite(e.c_eye))
        return oph_fail(ctx, OPH_ENUMERIC, "eye focus gives a non-finite effective focal length");

    const bool dev = flags & OPH_DEVICE_PTR;
    float2* A = static_cast<float2*>(oph_scratch(ctx, SLOT_WORK0, sizeof(float2) * n));
    float2* B = static_cast<float2*>(oph_scratch(ctx, SLOT_WORK1, sizeof(float2) * n));
    if (!A || !B) return OPH_ENOMEM;
    const double* dfr = fringe;
    double* dout = out;
    if (!dev) {
        double* buf = static_cast<double*>(oph_scratch(ctx, SLOT_IN, sizeof(double) * n));
        if (!buf) return OPH_ENOMEM;
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(buf, fringe, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
        dfr = buf;
        dout = buf;  // the fringe is consumed by recon_pre before the output is written
    }
    const unsigned grid = (unsigned)((n + 255) / 256);
    OPH_LAUNCH(ctx, "recon_pre", recon_pre, dim3(grid), dim3(256), 0, dfr, A, nx, ny);
    if ((rc = oph_fft_rows(ctx, nx, -1, A, ny, nx, 1.f))) return rc;
    if ((rc = oph_fft_cols(ctx, ny, -1, A, nx, nx, 1.f))) return rc;
    OPH_LAUNCH(ctx, "recon_eye", recon_eye, dim3(grid), dim3(256), 0, A, B, e);
    if ((rc = oph_fft_rows(ctx, nx, -1, B, ny, nx, 1.f))) return rc;
    if ((rc = oph_fft_cols(ctx, ny, -1, B, nx, nx, 1.f))) return rc;
    OPH_LAUNCH(ctx, "recon_mag", recon_mag, dim3(grid), dim3(256), 0, B, dout, n);
    if (!dev) OPH_HIP_CHECK(ctx, hipMemcpyAsync(out, dout, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    return oph_finish(ctx, flags);
}
