// pc_rs.hip -- exact-distance Rayleigh-Sommerfeld point-cloud field on gfx950.
//
// Replaces the KETI generator loop (K/ophgen_pointcloud/ophgen_pointcloud.cpp:
// 195-249) and its CUDA kernel (K/ophgen_pointcloud_gpu/ophgen_pointcloud_gpu.cpp:
// 391-415).  Same field, different machine mapping:
//   * a workgroup owns a 64x64 pixel tile, each thread a 4x4 pixel block, and
//     accumulates the complex field in registers (the reference does one
//     global read-modify-write per point per pixel);
//   * points are staged 256 at a time in LDS and read by broadcast;
//   * the phase k r needs ~1e6 rad of range at 1e-6 rad resolution, beyond
//     FP32: per (point, thread) the distance r0 to the block centre is formed
//     in FP64 and reduced to a fraction of a turn in FP64; per pixel the
//     offset r - r0 is the second-order expansion
//         (r - r0)/lambda = c1 q - c2 q^2,  q = r^2 - r0^2,
//         c1 = 1/(2 r0 lambda), c2 = 1/(8 r0^3 lambda)
//     in FP32 (q is a sum of a per-column and a per-row term), and sin/cos
//     are the hardware v_sin_f32/v_cos_f32, which take turns;
//   * a (thread, point) pair whose block is too close to the point for that
//     expansion (third-order term above 2e-8 turns, or |c1 q| above 16 turns)
//     takes an exact FP64 path per pixel;
//   * large point sets and small grids split the points over grid.z chunks;
//     a finalize kernel sums the chunks in a fixed order (deterministic),
//     applies the per-pixel tilt term in FP64 and writes the output format.
// The same kernel, with the amplitude divided by r per pixel and CWO's pixel
// grid, is cwoPLS::Huygens (CWO/cwoPLS.cpp:316-368):
//   U[i][j] = sum_n a_n exp(i (k r + ph + phi_n)) / r,
//   X_j = (j - nx/2) p, Y_i = (i - ny/2) p, r = |(X_j - x_n, Y_i - y_n, z_n + z)|;
// there 1/r = (1/r0)(1 + q/r0^2)^(-1/2) ~ (1/r0)(1 - q/(2 r0^2)) per pixel
// (relative error 3/8 (q/r0^2)^2 < 1e-8 wherever the phase expansion holds).
#include "oph_internal.h"

namespace {

constexpr int TPB_X = 16, TPB_Y = 16;     // threads per workgroup
constexpr int BX = 4, BY = 4;             // pixels per thread
constexpr int TILE_X = TPB_X * BX, TILE_Y = TPB_Y * BY;
constexpr int PBATCH = TPB_X * TPB_Y;     // points staged per batch

struct RsArgs {
    const double* x;
    const double* y;
    const double* z;
    const double* amp;
    const double* phase;
    int64_t n;
    int64_t chunk;          // points per grid.z chunk
    double sx, sy, sz, oz;  // scale, offset
    int nx, ny;
    double px, py;
    // pixel centre X(col) = col px + x_off, Y(row) = ydir row py + y_off
    double x_off, y_off, ydir;
    double ph0;             // global phase, turns
    double inv_lambda;
    float2* partial;        // [chunks][ny][nx]
};

template <bool HUYGENS>
__global__ __launch_bounds__(256) void rs_accumulate(RsArgs a) {
    __shared__ double s_x[PBATCH], s_y[PBATCH], s_zz[PBATCH], s_ph[PBATCH];
    __shared__ float s_a[PBATCH];

    const int tx = threadIdx.x, ty = threadIdx.y;
    const int tid = ty * TPB_X + tx;
    const int c0 = blockIdx.x * TILE_X + tx * BX;
    const int r0 = blockIdx.y * TILE_Y + ty * BY;
    // block centre (KETI: X = (col + 1/2) px - Lx/2, Y = Ly/2 - (row + 1/2) py)
    const double X0 = (c0 + 0.5 * (BX - 1)) * a.px + a.x_off;
    const double Y0 = a.ydir * (r0 + 0.5 * (BY - 1)) * a.py + a.y_off;
    float ex[BX], ey[BY];
#pragma unroll
    for (int c = 0; c < BX; ++c) ex[c] = (float)((c - 0.5 * (BX - 1)) * a.px);
#pragma unroll
    for (int r = 0; r < BY; ++r) ey[r] = (float)(a.ydir * (r - 0.5 * (BY - 1)) * a.py);
    // largest |offset| in the block (for the expansion bound)
    const double hmax = 0.5 * sqrt((double)(BX - 1) * (BX - 1) * a.px * a.px +
                                   (double)(BY - 1) * (BY - 1) * a.py * a.py);
    const float hmaxf = (float)hmax;

    float re[BY][BX], im[BY][BX];
#pragma unroll
    for (int r = 0; r < BY; ++r)
#pragma unroll
        for (int c = 0; c < BX; ++c) re[r][c] = im[r][c] = 0.f;

    const int64_t p_begin = (int64_t)blockIdx.z * a.chunk;
    int64_t p_end = p_begin + a.chunk;
    if (p_end > a.n) p_end = a.n;
    const float inv_l_f = (float)a.inv_lambda;

    for (int64_t base = p_begin; base < p_end; base += PBATCH) {
        __syncthreads();
        {
            const int64_t j = base + tid;
            double xs = 0.0, ys = 0.0, zz = 1.0, ph = 0.0;
            float am = 0.f;
            if (j < p_end) {
                xs = a.x[j] * a.sx;
                ys = a.y[j] * a.sy;
                const double zs = a.z[j] * a.sz + a.oz;
                zz = zs * zs;
                am = a.amp ? (float)a.amp[j] : 1.f;
                ph = (a.phase ? a.phase[j] * 0.15915494309189535 : 0.0) + a.ph0;  // rad -> turns
            }
            s_x[tid] = xs;
            s_y[tid] = ys;
            s_zz[tid] = zz;
            s_ph[tid] = ph;
            s_a[tid] = am;
        }
        __syncthreads();
        const int cnt = (int)((p_end - base) < PBATCH ? (p_end - base) : PBATCH);
        for (int j = 0; j < cnt; ++j) {
            const double dx0 = X0 - s_x[j];
            const double dy0 = Y0 - s_y[j];
            const double s0 = fma(dx0, dx0, fma(dy0, dy0, s_zz[j]));
            const double yr = rsqrt_f64(s0);            // 1/r0
            const double rr0 = s0 * yr;                   // r0
            const double t0 = fma(rr0, a.inv_lambda, s_ph[j]);
            const float f0 = (float)(t0 - rint(t0));
            const float am = s_a[j];
            const float dxf = (float)dx0, dyf = (float)dy0;
            const double c1d = 0.5 * a.inv_lambda * yr;
            const float c1 = (float)c1d;
            const float yrf = (float)yr;
            const float c2 = 0.25f * c1 * yrf * yrf;
            // expansion bound: u = q/r0^2 <= (2 h rho + h^2)/r0^2; the
            // third-order term r0 u^3/(16 lambda) and the first-order size
            // c1 q (FP32 resolution) decide the path
            // Huygens amplitude: am/r0 (1 - q/(2 r0^2)) per pixel
            const float amr = HUYGENS ? am * yrf : am;
            const float hq = HUYGENS ? -0.5f * yrf * yrf : 0.f;
            const float rhof = __builtin_sqrtf(fmaf(dxf, dxf, dyf * dyf));
            const float qmax = fmaf(2.f * hmaxf, rhof, hmaxf * hmaxf);
            const float u = qmax * yrf * yrf;
            const bool exact = ((float)rr0 * u * u * u * (0.0625f * inv_l_f) > 2e-8f) || (c1 * qmax > 16.f);
            if (!exact) {
                float A[BX], B[BY];
#pragma unroll
                for (int c = 0; c < BX; ++c) A[c] = fmaf(2.f * ex[c], dxf, ex[c] * ex[c]);
#pragma unroll
                for (int r = 0; r < BY; ++r) B[r] = fmaf(2.f * ey[r], dyf, ey[r] * ey[r]);
#pragma unroll
                for (int r = 0; r < BY; ++r)
#pragma unroll
                    for (int c = 0; c < BX; ++c) {
                        const float q = A[c] + B[r];
                        const float ph = fmaf(q, fmaf(-c2, q, c1), f0);
                        const float ap = HUYGENS ? amr * fmaf(hq, q, 1.f) : am;
                        re[r][c] = fmaf(ap, cos_turns(ph), re[r][c]);
                        im[r][c] = fmaf(ap, sin_turns(ph), im[r][c]);
                    }
            } else {
#pragma unroll
                for (int r = 0; r < BY; ++r)
#pragma unroll
                    for (int c = 0; c < BX; ++c) {
                        const double dx = dx0 + (c - 0.5 * (BX - 1)) * a.px;
                        const double dy = dy0 + a.ydir * (r - 0.5 * (BY - 1)) * a.py;
                        const double s = fma(dx, dx, fma(dy, dy, s_zz[j]));
                        const double ys = rsqrt_f64(s);
                        const double t = fma(s * ys, a.inv_lambda, s_ph[j]);
                        const float ph = (float)(t - rint(t));
                        const float ap = HUYGENS ? (float)(am * ys) : am;
                        re[r][c] = fmaf(ap, cos_turns(ph), re[r][c]);
                        im[r][c] = fmaf(ap, sin_turns(ph), im[r][c]);
                    }
            }
        }
    }
    float2* dst = a.partial + (size_t)blockIdx.z * a.nx * a.ny;
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        const int row = r0 + r;
        if (row >= a.ny) continue;
#pragma unroll
        for (int c = 0; c < BX; ++c) {
            const int col = c0 + c;
            if (col < a.nx) dst[(size_t)row * a.nx + col] = make_float2(re[r][c], im[r][c]);
        }
    }
}

struct FinArgs {
    const float2* partial;
    int chunks;
    int nx, ny;
    double px, py, half_lx, half_ly;
    double stx, sty, inv_lambda;  // sin(tilt)
    int tilt;
    void* out;
    int fmt;
    int accumulate;
};

// Sums the chunk partials in chunk order, applies exp(-i k (X sin tx + Y sin ty))
// and writes complex128 / complex64 / float Re.
__global__ __launch_bounds__(256) void rs_finalize(FinArgs f) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t npix = (int64_t)f.nx * f.ny;
    if (i >= npix) return;
    double re = 0.0, im = 0.0;
    for (int c = 0; c < f.chunks; ++c) {
        const float2 v = f.partial[(size_t)c * npix + i];
        re += v.x;
        im += v.y;
    }
    if (f.tilt) {
        const int row = (int)(i / f.nx), col = (int)(i % f.nx);
        const double X = (col + 0.5) * f.px - f.half_lx;
        const double Y = f.half_ly - (row + 0.5) * f.py;
        const double t = (X * f.stx + Y * f.sty) * f.inv_lambda;
        const float ft = (float)(t - rint(t));
        const double cs = cos_turns(ft), sn = -(double)sin_turns(ft);
        const double nr = re * cs - im * sn, ni = re * sn + im * cs;
        re = nr;
        im = ni;
    }
    if (f.fmt == OPH_OUT_C128) {
        double2* o = static_cast<double2*>(f.out) + i;
        if (f.accumulate) { double2 p = *o; re += p.x; im += p.y; }
        *o = make_double2(re, im);
    } else if (f.fmt == OPH_OUT_C64) {
        float2* o = static_cast<float2*>(f.out) + i;
        if (f.accumulate) { float2 p = *o; re += p.x; im += p.y; }
        *o = make_float2((float)re, (float)im);
    } else {
        float* o = static_cast<float*>(f.out) + i;
        float v = (float)re;
        if (f.accumulate) v += *o;
        *o = v;
    }
}

}  // namespace

namespace {

// Shared host side of oph_pc_rs and oph_pc_huygens: argument checks, point
// staging, the accumulate launch (KETI or CWO pixel grid, with or without
// 1/r) and the finalize launch.
struct RsJob {
    int nx, ny;
    double px, py, lambda;
    double scale[3], offset_z;
    double tilt_deg[2];
    bool cwo_grid, huygens;
    double ph0;  // global phase (radians)
};

int rs_run(oph_ctx* ctx, const RsJob& J, const oph_points* pts, void* out, int flags) {
    if (!pts || !out) return oph_fail(ctx, OPH_EINVAL, "null argument");
    if (J.nx <= 0 || J.ny <= 0 || J.nx > 65536 || J.ny > 65536)
        return oph_fail(ctx, OPH_EINVAL, "image size %dx%d out of range", J.nx, J.ny);
    if (!(J.px > 0) || !(J.py > 0) || !(J.lambda > 0))
        return oph_fail(ctx, OPH_EINVAL, "pitch and wavelength must be positive");
    if (pts->n < 0 || (pts->n > 0 && (!pts->x || !pts->y || !pts->z)))
        return oph_fail(ctx, OPH_EINVAL, "bad point arrays");
    const int fmt = flags & OPH_OUT_MASK;
    if (fmt != OPH_OUT_C128 && fmt != OPH_OUT_C64 && fmt != OPH_OUT_FRINGE)
        return oph_fail(ctx, OPH_EINVAL, "bad output format flag");
    if ((flags & OPH_ASYNC) && !(flags & OPH_DEVICE_PTR))
        return oph_fail(ctx, OPH_EINVAL, "OPH_ASYNC needs device pointers");

    const int nx = J.nx, ny = J.ny;
    const int64_t n = pts->n;
    const size_t npix = (size_t)nx * ny;
    const size_t pbytes = (size_t)(n > 0 ? n : 1) * sizeof(double);

    RsArgs a{};
    a.n = n;
    if (!(flags & OPH_DEVICE_PTR) && !oph_scratch(ctx, SLOT_PTS, 5 * pbytes)) return OPH_ENOMEM;
    const int sflags = n > 0 ? flags : (flags | OPH_DEVICE_PTR);  // nothing to copy when n == 0
    a.x = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->x, pbytes, sflags, 0 * pbytes));
    a.y = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->y, pbytes, sflags, 1 * pbytes));
    a.z = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->z, pbytes, sflags, 2 * pbytes));
    a.amp = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->amp, pbytes, sflags, 3 * pbytes));
    a.phase = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->phase, pbytes, sflags, 4 * pbytes));
    if (n > 0 && (!a.x || !a.y || !a.z)) return OPH_ENOMEM;
    if (n > 0 && ((pts->amp && !a.amp) || (pts->phase && !a.phase))) return OPH_ENOMEM;

    a.sx = J.scale[0];
    a.sy = J.scale[1];
    a.sz = J.scale[2];
    a.oz = J.offset_z;
    a.nx = nx;
    a.ny = ny;
    a.px = J.px;
    a.py = J.py;
    const double half_lx = J.px * nx / 2.0, half_ly = J.py * ny / 2.0;
    if (J.cwo_grid) {
        // cwoPLS: X = (col - nx/2) p, Y = (row - ny/2) p (CWO/cwoPLS.cpp:350-351)
        a.x_off = -(double)(nx / 2) * J.px;
        a.y_off = -(double)(ny / 2) * J.py;
        a.ydir = 1.0;
    } else {
        // KETI: X = (col + 1/2) p - Lx/2, Y = Ly/2 - (row + 1/2) p
        a.x_off = 0.5 * J.px - half_lx;
        a.y_off = half_ly - 0.5 * J.py;
        a.ydir = -1.0;
    }
    a.ph0 = J.ph0 * 0.15915494309189535;
    a.inv_lambda = 1.0 / J.lambda;

    const int gx = (nx + TILE_X - 1) / TILE_X, gy = (ny + TILE_Y - 1) / TILE_Y;
    const int64_t tiles = (int64_t)gx * gy;
    // split points so that the grid holds >= 4 workgroups per CU, keeping
    // >= 1024 points per chunk
    int64_t chunks = (4LL * ctx->cu_count + tiles - 1) / tiles;
    const int64_t max_by_pts = (n + 1023) / 1024;
    if (chunks > max_by_pts) chunks = max_by_pts;
    if (chunks > 16) chunks = 16;
    if (chunks < 1) chunks = 1;
    a.chunk = (n + chunks - 1) / chunks;
    if (a.chunk < 1) a.chunk = 1;
    a.partial = static_cast<float2*>(oph_scratch(ctx, SLOT_WORK0, chunks * npix * sizeof(float2)));
    if (!a.partial) return OPH_ENOMEM;

    if (J.huygens)
        OPH_LAUNCH(ctx, "huygens_accumulate", rs_accumulate<true>, dim3(gx, gy, (unsigned)chunks),
                   dim3(TPB_X, TPB_Y), 0, a);
    else
        OPH_LAUNCH(ctx, "rs_accumulate", rs_accumulate<false>, dim3(gx, gy, (unsigned)chunks),
                   dim3(TPB_X, TPB_Y), 0, a);

    const size_t obytes = npix * (fmt == OPH_OUT_C128 ? 16 : fmt == OPH_OUT_C64 ? 8 : 4);
    void* dout = out;
    if (!(flags & OPH_DEVICE_PTR)) {
        dout = oph_scratch(ctx, SLOT_OUT, obytes);
        if (!dout) return OPH_ENOMEM;
        if (flags & OPH_ACCUMULATE)
            OPH_HIP_CHECK(ctx, hipMemcpyAsync(dout, out, obytes, hipMemcpyHostToDevice, ctx->stream));
    }
    FinArgs f{};
    f.partial = a.partial;
    f.chunks = (int)chunks;
    f.nx = nx;
    f.ny = ny;
    f.px = a.px;
    f.py = a.py;
    f.half_lx = half_lx;
    f.half_ly = half_ly;
    f.stx = sin(J.tilt_deg[0] * (M_PI / 180.0));
    f.sty = sin(J.tilt_deg[1] * (M_PI / 180.0));
    f.inv_lambda = a.inv_lambda;
    f.tilt = (f.stx != 0.0 || f.sty != 0.0);
    f.out = dout;
    f.fmt = fmt;
    f.accumulate = (flags & OPH_ACCUMULATE) ? 1 : 0;
    OPH_LAUNCH(ctx, "rs_finalize", rs_finalize, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, f);
    if (!(flags & OPH_DEVICE_PTR))
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(out, dout, obytes, hipMemcpyDeviceToHost, ctx->stream));
    return oph_finish(ctx, flags);
}

}  // namespace

extern "C" int oph_pc_rs(oph_ctx* ctx, const oph_spec* spec, const oph_points* pts, void* out,
                         int flags) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!spec) return oph_fail(ctx, OPH_EINVAL, "null argument");
    RsJob J{};
    J.nx = spec->nx;
    J.ny = spec->ny;
    J.px = spec->pitch_x;
    J.py = spec->pitch_y;
    J.lambda = spec->lambda;
    for (int k = 0; k < 3; ++k) J.scale[k] = spec->scale[k];
    J.offset_z = spec->offset_z;
    J.tilt_deg[0] = spec->tilt_deg[0];
    J.tilt_deg[1] = spec->tilt_deg[1];
    return rs_run(ctx, J, pts, out, flags);
}

extern "C" int oph_pc_huygens(oph_ctx* ctx, int nx, int ny, double pitch, double lambda, double z, double ph,
                              const oph_points* pts, void* out, int flags) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if ((flags & OPH_OUT_MASK) == OPH_OUT_FRINGE)
        return oph_fail(ctx, OPH_EINVAL, "oph_pc_huygens writes complex output only");
    RsJob J{};
    J.nx = nx;
    J.ny = ny;
    J.px = J.py = pitch;
    J.lambda = lambda;
    J.scale[0] = J.scale[1] = J.scale[2] = 1.0;
    J.offset_z = z;
    J.cwo_grid = true;
    J.huygens = true;
    J.ph0 = ph;
    return rs_run(ctx, J, pts, out, flags);
}
