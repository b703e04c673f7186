/*
 * oph_oracle.c -- CPU restatement of OpenHolo's hologram-generation algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file: it is the CPU restatement that tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg compare the HIP path against.
 *
 * Every routine follows a reference file:line (paths relative to the reference
 * tree; K/ = OpenHolo_KETI/VS/Openholo_Generation_PointCloud_Library_beta_v1.0,
 * DM/ = OpenHolo_DepthMap, W/ = OpenHolo_WRP/openwrp_console_ver 1.0,
 * CWO/ = W/cwo/src).  Arithmetic is FP64 unless the routine says otherwise.
 * Loops over pixels are split over OpenMP threads by output row only, so each
 * output element is produced by one thread in the reference's summation order
 * (the reference's own OpenMP loops race on the output; this one does not).
 *
 * How each function is pinned (the tests that hold it; tests/test_oracle_*.py):
 *   orc_keti_fringe_f32   bit-exact vs the compiled reference source (oracle/_ref)
 *   orc_rs                single-point closed form, linearity, translation
 *                         equivariance, Re() vs the FP32 reference fixtures
 *   orc_fresnel           brute-force long-double per-pair sum on tiny cases,
 *                         rank-1 separable closed form of one plane, and the
 *                         paraxial limit of orc_rs (same points, pixel origins
 *                         aligned: phase sign, carrier and grid conventions)
 *   orc_fft               numpy.fft on random inputs, O(N^2) DFT, Parseval
 *   orc_asm_transfer      values against exp(i 2 pi z sqrt(..)), band limit on
 *                         both axes against Matsushima's closed-form cutoff
 *                         |f| <= 1/(lambda sqrt((2 df z)^2 + 1))
 *   orc_asm               plane-wave eigenfunction (with and without the band
 *                         limit), forward/back identity, evanescent removal,
 *                         spherical wave carried from z1 to z1 + dz, paraxial
 *                         limit against the Fresnel transfer function, padded
 *                         result = crop(IFFT(orc_asm_transfer . FFT(embed)))
 *   orc_huygens           long-double brute force on tiny cases; single point
 *                         against orc_rs (|U| r = a, same phase) with the grids
 *                         aligned (half pixel, row flip)
 *   orc_wrp               single-point window vs the closed-form spherical wave,
 *                         window size and clipping, linearity
 *   orc_dm_encode         explicit inverse-DFT sum over the kept spectrum rows
 *                         with the shift phase (bins on both sides of both crop
 *                         edges, complex amplitudes, padded and unpadded grids)
 *   orc_depthmap          one-pixel scene: spectrum and fringe against closed
 *                         forms (with and without the field lens), layer
 *                         linearity, depth-index modes, level drift
 *   orc_normalize_u8      bit-exact vs the compiled reference ophNormalize
 *   orc_scale_u8          hand-computed values (min, max, interior, constant)
 *   orc_dm_reconstruct    closed form: with the eye focused at distance 0 the
 *                         eye-lens chirp cancels the field-lens chirp and the
 *                         result is N |fringe| reflected about the centre
 *                         (numpy FFT check of two shifted transforms); the
 *                         eye-centre shift moves the image by the expected
 *                         number of pixels; pupil mask energy monotonicity
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#define TWO_PI (2.0 * M_PI)

typedef struct { double re, im; } c128;

static inline c128 cmul(c128 a, c128 b) {
    c128 r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline c128 cexpi(double ph) { c128 r = {cos(ph), sin(ph)}; return r; }

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* ------------------------------------------------------------------------ */
/* Point-cloud Rayleigh-Sommerfeld field (KETI)                              */
/* ------------------------------------------------------------------------ */

/* Spec as the KETI OphSpec fields the generator reads
 * (K/ophgen_pointcloud/ophgen_pointcloud.h:45-70), carried in double. */
typedef struct {
    int nx, ny;
    double pitch_x, pitch_y;
    double lambda;
    double scale[3];
    double offset_z;
    double tilt_deg[2];
} orc_spec;

/*
 * Complex RS field, FP64.  K/ophgen_pointcloud/ophgen_pointcloud.cpp:195-249:
 *   point  x = x*sx, y = y*sy, z = z*sz + offset             (:224-227)
 *   pixel  X = (col+1/2)p - Lx/2,  Y = Ly/2 - (row+1/2)p      (:231,235)
 *   r = sqrt((X-x)^2 + (Y-y)^2 + z^2),  k = 2 pi / lambda     (:205,237)
 *   phase  = k r - k X sin(tx) - k Y sin(ty)                  (:238)
 * The reference stores a*cos(phase) (:239-241); the complex form returns
 *   U[row][col] = sum_n a_n exp(i (phase + phi_n))
 * so that Re(U) at phi = 0 is the reference fringe.  phi may be NULL (the
 * reference parses the point phase but never uses it, :224-241).
 */
void orc_rs(const orc_spec* s, int64_t n, const double* x, const double* y,
            const double* z, const double* amp, const double* phi,
            c128* out, int nthreads) {
    set_threads(nthreads);
    const int nx = s->nx, ny = s->ny;
    const double k = TWO_PI / s->lambda;
    const double Lx = s->pitch_x * nx, Ly = s->pitch_y * ny;
    const double stx = sin(s->tilt_deg[0] * (M_PI / 180.0));
    const double sty = sin(s->tilt_deg[1] * (M_PI / 180.0));
    double* px = (double*)malloc(sizeof(double) * (size_t)n * 5);
    double *py = px + n, *pz = py + n, *pa = pz + n, *pp = pa + n;
    for (int64_t j = 0; j < n; ++j) {
        px[j] = x[j] * s->scale[0];
        py[j] = y[j] * s->scale[1];
        pz[j] = z[j] * s->scale[2] + s->offset_z;
        pa[j] = amp[j];
        pp[j] = phi ? phi[j] : 0.0;
    }
    int row;
#pragma omp parallel for schedule(dynamic, 1)
    for (row = 0; row < ny; ++row) {
        const double Y = Ly / 2 - (row + 0.5) * s->pitch_y;
        for (int col = 0; col < nx; ++col) {
            const double X = (col + 0.5) * s->pitch_x - Lx / 2;
            double re = 0.0, im = 0.0;
            for (int64_t j = 0; j < n; ++j) {
                const double dx = X - px[j], dy = Y - py[j];
                const double r = sqrt(dx * dx + dy * dy + pz[j] * pz[j]);
                /* reduce k r in revolutions before the trig call: r/lambda is
                 * ~2e5 turns here, so fmod on the turn count keeps the FP64
                 * argument small and exact to ~1e-11 rad */
                double t = r / s->lambda;
                t -= floor(t);
                const double ph = TWO_PI * t + pp[j];
                re += pa[j] * cos(ph);
                im += pa[j] * sin(ph);
            }
            /* tilt term is per pixel (independent of the point), :238 */
            double tt = (X * stx + Y * sty) / s->lambda;
            tt -= floor(tt);
            const c128 tilt = cexpi(-TWO_PI * tt);
            c128 u = {re, im};
            out[(size_t)row * nx + col] = cmul(u, tilt);
        }
    }
    (void)k;
    free(px);
}

/* KETI OphSpec, float fields, field order of ophgen_pointcloud.h:45-70. */
typedef struct {
    float scale[3];
    float offset_depth;
    float pitch[2];
    int nx, ny;
    float lambda;
    float tilt[2];
} orc_keti_spec_f32;

/*
 * Reference-faithful FP32 fringe: the reference's own float arithmetic of
 * K/ophgen_pointcloud/ophgen_pointcloud.cpp:195-249 with the point loop
 * innermost, which keeps each pixel's summation order n = 0..P-1 (the
 * single-thread order of :223-243).  dst is accumulated into (+=), as in
 * the reference; the caller zeroes it (example_cpu.cpp:44).
 * Built with -ffp-contract=off so no FMA changes the rounding.
 */
void orc_keti_fringe_f32(const orc_keti_spec_f32* c, int n, const float* xyz,
                         const float* amp, float* dst, int nthreads) {
    set_threads(nthreads);
    const int Nx = c->nx, Ny = c->ny;
    const float thetaX = c->tilt[0] * (M_PI / 180.f);
    const float thetaY = c->tilt[1] * (M_PI / 180.f);
    const float k = (2.f * M_PI) / c->lambda;
    const float pixel_x = c->pitch[0], pixel_y = c->pitch[1];
    const float Length_x = pixel_x * Nx, Length_y = pixel_y * Ny;
    int row;
#pragma omp parallel for schedule(dynamic, 1)
    for (row = 0; row < Ny; ++row) {
        const float SLM_y = (Length_y / 2) - ((float)row + 0.5f) * pixel_y;
        for (int col = 0; col < Nx; ++col) {
            const float SLM_x = ((float)col + 0.5f) * pixel_x - (Length_x / 2);
            float acc = dst[col + row * Nx];
            for (int j = 0; j < n; ++j) {
                const float x = xyz[3 * j + 0] * c->scale[0];
                const float y = xyz[3 * j + 1] * c->scale[1];
                const float z = xyz[3 * j + 2] * c->scale[2] + c->offset_depth;
                const float r = sqrtf((SLM_x - x) * (SLM_x - x) + (SLM_y - y) * (SLM_y - y) + z * z);
                const float phi = k * r - k * SLM_x * sinf(thetaX) - k * SLM_y * sinf(thetaY);
                acc += amp[j] * cosf(phi);
            }
            dst[col + row * Nx] = acc;
        }
    }
}

/*
 * 8-bit normalisation, K/oph_common/oph_common.cpp:92-116: global min/max,
 * vertical flip (:111), (v-min)/(max-min)*255 + 0.5 truncated to u8 (:113).
 * Float arithmetic as the reference.
 */
void orc_normalize_u8(const float* src, unsigned char* dst, int nx, int ny) {
    float mn = src[0], mx = src[0];
    for (int i = 0; i < nx * ny; ++i) {
        if (src[i] < mn) mn = src[i];
        if (src[i] > mx) mx = src[i];
    }
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x)
            dst[x + (ny - y - 1) * nx] =
                (unsigned char)(((src[x + y * nx] - mn) / (mx - mn)) * 255 + 0.5);
}

/* ------------------------------------------------------------------------ */
/* Fresnel point sum (CWO cwoPLS::Fresnel)                                   */
/* ------------------------------------------------------------------------ */

/*
 * Direct Fresnel sum, FP64.  CWO/cwoPLS.cpp:370-422:
 *   pixel  X = (j - Nx/2) p,  Y = (i - Ny/2) p   (:400-401, zero offsets)
 *   phase  = k ((X-x_n)^2 + (Y-y_n)^2) / (2 z_n)  (:404-406)
 * extended as the depth-plane point CGH needs it with the point amplitude,
 * its initial phase and the plane carrier exp(i k z_n):
 *   U[i][j] = sum_n a_n exp(i (k z_n + phi_n + k rho^2/(2 z_n)))
 * Only pixels listed in (rows, cols) are evaluated when rows != NULL
 * (sampled-pixel oracle for full-size grids); out has npix entries then.
 */
void orc_fresnel(int nx, int ny, double pitch, double lambda, int64_t n,
                 const double* x, const double* y, const double* z,
                 const double* amp, const double* phi, const int* rows,
                 const int* cols, int64_t npix, c128* out, int nthreads) {
    set_threads(nthreads);
    const double inv_l = 1.0 / lambda;
    const int64_t total = rows ? npix : (int64_t)nx * ny;
    int64_t q;
#pragma omp parallel for schedule(dynamic, 64)
    for (q = 0; q < total; ++q) {
        const int i = rows ? rows[q] : (int)(q / nx);
        const int j = rows ? cols[q] : (int)(q % nx);
        const double X = (j - nx / 2) * pitch, Y = (i - ny / 2) * pitch;
        double re = 0.0, im = 0.0;
        for (int64_t m = 0; m < n; ++m) {
            const double dx = X - x[m], dy = Y - y[m];
            /* turns = z/lambda + rho^2/(2 z lambda); reduce each part */
            double t1 = z[m] * inv_l;
            t1 -= floor(t1);
            double t2 = (dx * dx + dy * dy) / (2.0 * z[m]) * inv_l;
            t2 -= floor(t2);
            const double ph = TWO_PI * (t1 + t2) + (phi ? phi[m] : 0.0);
            re += amp[m] * cos(ph);
            im += amp[m] * sin(ph);
        }
        out[q].re = re;
        out[q].im = im;
    }
}

/*
 * Huygens point sum, FP64, CWO/cwoPLS.cpp:316-368 (cwoPLS::Huygens) with the
 * point amplitude and initial phase (NULL: 1 and 0):
 *   pixel  X = (j - Nx/2) p,  Y = (i - Ny/2) p     (:350-351, zero offsets)
 *   dz = z_n + z                                  (:352)
 *   U[i][j] = sum_n a_n exp(i (k r + ph + phi_n)) / r,  r = |(X - x, Y - y, dz)|
 */
void orc_huygens(int nx, int ny, double pitch, double lambda, double z, double ph, int64_t n,
                 const double* x, const double* y, const double* zz, const double* amp,
                 const double* phi, c128* out, int nthreads) {
    set_threads(nthreads);
    int i;
#pragma omp parallel for schedule(dynamic, 1)
    for (i = 0; i < ny; ++i) {
        const double Y = (i - ny / 2) * pitch;
        for (int j = 0; j < nx; ++j) {
            const double X = (j - nx / 2) * pitch;
            double re = 0.0, im = 0.0;
            for (int64_t m = 0; m < n; ++m) {
                const double dx = X - x[m], dy = Y - y[m], dz = zz[m] + z;
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                double t = r / lambda;
                t -= floor(t);
                const double p = TWO_PI * t + ph + (phi ? phi[m] : 0.0);
                const double a = (amp ? amp[m] : 1.0) / r;
                re += a * cos(p);
                im += a * sin(p);
            }
            out[(size_t)i * nx + j].re = re;
            out[(size_t)i * nx + j].im = im;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Wavefront-recording plane (OPHWRP::SingleWRP)                             */
/* ------------------------------------------------------------------------ */

/*
 * WRP accumulation, FP64.  W/openwrp/WRP.cpp:75-142 with CWO::AddPixel
 * (CWO/cwo.cpp:4298-4309):
 *   dz = z_wrp - z_n                                   (:110)
 *   centre tx = (int)(x_n/p) + Nx/2, ty likewise        (:114-115, truncation)
 *   window wx, wy in [-w, w)                            (:119-120)
 *   r = sign(dz) sqrt(dx^2 + dy^2 + dz^2), dx = wx p     (:122-128)
 *   wrp[ty+wy][tx+wx] += a_n exp(i k r) / (r + 0.05)     (:131-136), in bounds
 * Half-width w: mode 0 = grating-angle radius of CWO/cwoPLS.cpp:613
 *   w = (int)(|dz| tan(asin(lambda/(2p))) / p);
 * mode 1 = W/openwrp/WRP.cpp:111, w = (int)|lambda dz/(2p^2) + 0.5|*2 - 1.
 * mode 2 = cwoPLS::WRP (CWO/cwoPLS.cpp:522-591): w = (int)(|dz| lambda/(2p)/p)
 *   (:557), Fresnel phase r = dz + (dx^2 + dy^2)/(2 dz) (:573), amplitude a
 *   (the reference's unit amplitude when amp is NULL; no 1/(r + 0.05)).
 * mode 3 = the MATLAB statement (OpenHolo_WRP/WRP_matlab_ver 1.0/
 *   Hologram_WRP.m:19-36, fn_wrp.m): N = round(|lambda dz/p^2/2| + 0.5) 2 - 1
 *   as the FULL width, a square window wx, wy in [-(N-1)/2, (N-1)/2]
 *   centred at round(x/p) + (Nx-1)/2 - 1 (the script's 1-based
 *   round(x/p) + (res-1)/2, made 0-based; its matrix holds x along rows,
 *   here along columns), sub-hologram exp(i phase_n) exp(i k r)/r with
 *   r = sign(dz) sqrt(dx^2 + dy^2 + dz^2) (:29-30; the script's random
 *   phase rand*2*pi is the caller's phase_n); points with dz = 0 (r = 0 at
 *   the centre) contribute nothing.
 * amp may be NULL (unit amplitude, as the reference's cwoObjPoint); phase
 * (radians) is used by mode 3 only and may be NULL (zero).
 * The reference adds points in parallel with a racing AddPixel; here each
 * WRP pixel sums its contributions in point order n = 0..P-1.
 */
void orc_wrp(int nx, int ny, double pitch, double lambda, double z_wrp, int mode,
             int64_t n, const double* x, const double* y, const double* z,
             const double* amp, const double* phase, c128* wrp, int nthreads) {
    set_threads(nthreads);
    const int nxh = nx >> 1, nyh = ny >> 1;
    int* tx = (int*)malloc(sizeof(int) * (size_t)n * 4);
    int *ty = tx + n, *tw = ty + n, *tlo = tw + n;
    const double ga = tan(asin(lambda / (2.0 * pitch)));
    for (int64_t m = 0; m < n; ++m) {
        const double dz = z_wrp - z[m];
        tx[m] = (int)(x[m] / pitch) + nxh;
        ty[m] = (int)(y[m] / pitch) + nyh;
        if (mode == 0)
            tw[m] = (int)(fabs(dz) * ga / pitch);
        else if (mode == 1)
            tw[m] = (int)fabs(lambda * dz / pitch / pitch / 2 + 0.5) * 2 - 1;
        else if (mode == 2)
            tw[m] = (int)(fabs(dz) * (lambda / (2.0 * pitch)) / pitch);
        tlo[m] = -tw[m];
        if (mode == 3) {
            const int N = (int)round(fabs(lambda * dz / (pitch * pitch) / 2) + 0.5) * 2 - 1;
            tx[m] = (int)round(x[m] / pitch) + (nx - 1) / 2 - 1;
            ty[m] = (int)round(y[m] / pitch) + (ny - 1) / 2 - 1;
            tlo[m] = -(N - 1) / 2;
            tw[m] = dz == 0.0 ? -(N - 1) / 2 : (N - 1) / 2 + 1;  /* window [tlo, tw) */
        }
    }
    memset(wrp, 0, sizeof(c128) * (size_t)nx * ny);
    int row;
#pragma omp parallel for schedule(dynamic, 4)
    for (row = 0; row < ny; ++row) {
        for (int64_t m = 0; m < n; ++m) {
            const int w = tw[m];
            const int wy = row - ty[m];
            if (wy < tlo[m] || wy >= w) continue;
            const double dz = z_wrp - z[m];
            const double sgn = dz > 0.0 ? 1.0 : -1.0;
            const double a = amp ? amp[m] : 1.0;
            const double ph = (mode == 3 && phase) ? phase[m] / TWO_PI : 0.0;
            for (int wx = tlo[m]; wx < w; ++wx) {
                const int col = tx[m] + wx;
                if (col < 0 || col >= nx) continue;
                const double dx = wx * pitch, dy = wy * pitch;
                const double r = mode == 2 ? dz + (dx * dx + dy * dy) / (2.0 * dz)
                                           : sgn * sqrt(dx * dx + dy * dy + dz * dz);
                double t = r / lambda + ph;
                t -= floor(t);
                const double inv = mode == 2 ? a : mode == 3 ? a / r : a / (r + 0.05);
                c128* o = &wrp[(size_t)row * nx + col];
                o->re += cos(TWO_PI * t) * inv;
                o->im += sin(TWO_PI * t) * inv;
            }
        }
    }
    free(tx);
}

/* ------------------------------------------------------------------------ */
/* FFT (own FP64 Stockham radix-2; the reference uses FFTW / cuFFT Z2Z)      */
/* ------------------------------------------------------------------------ */

/* In-place 1-D DFT of length n (power of two), sign -1 forward (FFTW_FORWARD),
 * +1 backward, unnormalised, as FFTW's convention.  Stockham autosort radix-2
 * with a twiddle table tw[m] = exp(sign i 2 pi m / n), m < n/2, computed in
 * FP64 by the caller. */
static void fft1d(c128* a, c128* work, int n, const c128* tw) {
    c128* src = a;
    c128* dst = work;
    const int half = n >> 1;
    for (int ns = 1; ns < n; ns <<= 1) {
        const int step = half / ns;
        for (int j = 0; j < half; ++j) {
            const int k = j % ns;
            const c128 u = src[j], v = cmul(src[j + half], tw[k * step]);
            const int d = (j / ns) * ns * 2 + k;
            dst[d].re = u.re + v.re; dst[d].im = u.im + v.im;
            dst[d + ns].re = u.re - v.re; dst[d + ns].im = u.im - v.im;
        }
        c128* t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, sizeof(c128) * n);
}

/* In-place 1-D DFT of any length n, sign and normalisation as fft1d:
 * Stockham autosort over the prime factors of n (smallest first), each
 * radix-p butterfly a direct p-point DFT, with tw[m] = exp(sign i 2 pi m / n)
 * for m < n.  Stage with radix p after radices of product ns: inputs
 * src[j + r n/p] (r < p) times exp(sign i 2 pi r k/(ns p)), k = j mod ns,
 * then the p-point DFT, outputs to dst[(j/ns) ns p + k + q ns] (q < p) --
 * the radix-2 loop above is the case p = 2. */
static void fft1d_mixed(c128* a, c128* work, int n, const c128* tw) {
    c128* src = a;
    c128* dst = work;
    int ns = 1, rest = n;
    while (rest > 1) {
        int p = 2;
        while (rest % p) ++p;
        const int nb = n / p;
        const size_t step = (size_t)n / ((size_t)ns * p);
        c128* v = (c128*)malloc(sizeof(c128) * 2 * (size_t)p);
        c128* o = v + p;
        for (int j = 0; j < nb; ++j) {
            const int k = j % ns;
            for (int r = 0; r < p; ++r) {
                const c128 x = src[j + (size_t)r * nb];
                v[r] = r ? cmul(x, tw[(size_t)r * k * step]) : x;
            }
            for (int q = 0; q < p; ++q) {
                c128 acc = {0.0, 0.0};
                for (int r = 0; r < p; ++r) {
                    const c128 t = cmul(v[r], tw[((size_t)r * q % p) * nb]);
                    acc.re += t.re;
                    acc.im += t.im;
                }
                o[q] = acc;
            }
            const size_t base = (size_t)(j / ns) * ns * p + k;
            for (int q = 0; q < p; ++q) dst[base + (size_t)q * ns] = o[q];
        }
        free(v);
        c128* t = src; src = dst; dst = t;
        ns *= p;
        rest /= p;
    }
    if (src != a) memcpy(a, src, sizeof(c128) * n);
}

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* exp(sign i 2 pi m / n) for m < n/2 (power of two, fft1d) or m < n (fft1d_mixed) */
static c128* twiddles(int n, int sign) {
    const int cnt = is_pow2(n) ? n / 2 + 1 : n;
    c128* tw = (c128*)malloc(sizeof(c128) * cnt);
    for (int m = 0; m < cnt; ++m) tw[m] = cexpi(sign * TWO_PI * (double)m / (double)n);
    return tw;
}

static void fft1d_any(c128* a, c128* work, int n, const c128* tw) {
    if (is_pow2(n)) fft1d(a, work, n, tw);
    else fft1d_mixed(a, work, n, tw);
}

/* 2-D DFT of a row-major ny x nx array, in place; sign as fft1d;
 * normalize != 0 divides by nx*ny. */
void orc_fft2d(c128* a, int nx, int ny, int sign, int normalize, int nthreads) {
    set_threads(nthreads);
    c128* twx = twiddles(nx, sign);
    c128* twy = twiddles(ny, sign);
    int row;
#pragma omp parallel for
    for (row = 0; row < ny; ++row) {
        c128* w = (c128*)malloc(sizeof(c128) * nx);
        fft1d_any(a + (size_t)row * nx, w, nx, twx);
        free(w);
    }
    int cb;
    const int B = 8;  /* columns gathered together for cache locality */
#pragma omp parallel for
    for (cb = 0; cb < (nx + B - 1) / B; ++cb) {
        c128* line = (c128*)malloc(sizeof(c128) * ny * 2);
        const double s = normalize ? 1.0 / ((double)nx * ny) : 1.0;
        for (int col = cb * B; col < nx && col < cb * B + B; ++col) {
            for (int r = 0; r < ny; ++r) line[r] = a[(size_t)r * nx + col];
            fft1d_any(line, line + ny, ny, twy);
            for (int r = 0; r < ny; ++r) {
                a[(size_t)r * nx + col].re = line[r].re * s;
                a[(size_t)r * nx + col].im = line[r].im * s;
            }
        }
        free(line);
    }
    free(twx);
    free(twy);
}

/* ------------------------------------------------------------------------ */
/* Angular-spectrum propagation with zero padding                            */
/* ------------------------------------------------------------------------ */

/*
 * Transfer value of the angular-spectrum method at FFT-order bin (r, c) of an
 * Npy x Npx grid (fx = kx/(Npx px), kx in [-Npx/2, Npx/2), fy likewise):
 *   H = exp(i 2 pi z sqrt(1/lambda^2 - fx^2 - fy^2)), 0 if evanescent
 * (CWO_ANGULAR, CWO/cwo_lib.h:66-101; fn_FresnelPropagation_as.m:24).
 * band_limit zeroes H where |fx z / sqrt(..)| > 1/(2 dfx) or
 * |fy z / sqrt(..)| > 1/(2 dfy), dfx = 1/(Npx px) the grid's frequency step
 * (fn_FresnelPropagation_as.m:33-38 with sx = sy = 0: Matsushima's band limit).
 */
static c128 asm_h(int r, int c, int Npx, int Npy, double px, double py, double lambda,
                  double z, int band_limit) {
    const c128 zero = {0.0, 0.0};
    const double dfx = 1.0 / (Npx * px), dfy = 1.0 / (Npy * py);
    const int ky = r < Npy / 2 ? r : r - Npy;
    const int kx = c < Npx / 2 ? c : c - Npx;
    const double fx = kx * dfx, fy = ky * dfy;
    const double s = 1.0 / (lambda * lambda) - fx * fx - fy * fy;
    if (s <= 0.0) return zero;
    const double sq = sqrt(s);
    if (band_limit) {
        const double fla = fabs(fx * z / sq), flb = fabs(fy * z / sq);
        if (fla > 1.0 / (2.0 * dfx) || flb > 1.0 / (2.0 * dfy)) return zero;
    }
    double t = z * sq;  /* turns */
    t -= floor(t);
    return cexpi(TWO_PI * t);
}

/* The transfer of orc_asm on the whole Npy x Npx grid, FFT order. */
void orc_asm_transfer(int Npx, int Npy, double px, double py, double lambda, double z,
                      int band_limit, c128* H) {
    for (int r = 0; r < Npy; ++r)
        for (int c = 0; c < Npx; ++c)
            H[(size_t)r * Npx + c] = asm_h(r, c, Npx, Npy, px, py, lambda, z, band_limit);
}

/*
 * ASM propagation of an ny x nx field over distance z, FP64.
 * Plumbing of CWO::Diffract (CWO/cwo.cpp:2198-2278): zero-pad into a
 * (pad*ny) x (pad*nx) grid at offset (Npx/2 - nx/2, Npy/2 - ny/2)
 * (:2235-2239), transform, crop the same window back (:2258-2263).
 * Transfer asm_h above on the padded grid (band_limit as there).
 * The inverse transform is normalised by 1/(Npx Npy) (cwo.cpp:2930).
 * (fn_FresnelPropagation_as.m evaluates its transfer on the grid
 * ((1:N) - N/2) dal, which after its fftshift sits one bin above each
 * sample's frequency; the FFT-order grid here is CWO's and the exact one.)
 */
void orc_asm(c128* field, int nx, int ny, double px, double py, double lambda,
             double z, int pad, int band_limit, int nthreads) {
    const int Npx = nx * pad, Npy = ny * pad;
    const int ox = Npx / 2 - nx / 2, oy = Npy / 2 - ny / 2;
    c128* g = (c128*)calloc((size_t)Npx * Npy, sizeof(c128));
    for (int r = 0; r < ny; ++r)
        memcpy(g + (size_t)(r + oy) * Npx + ox, field + (size_t)r * nx, sizeof(c128) * nx);
    orc_fft2d(g, Npx, Npy, -1, 0, nthreads);
    for (int r = 0; r < Npy; ++r)
        for (int c = 0; c < Npx; ++c) {
            c128* e = &g[(size_t)r * Npx + c];
            *e = cmul(*e, asm_h(r, c, Npx, Npy, px, py, lambda, z, band_limit));
        }
    orc_fft2d(g, Npx, Npy, +1, 1, nthreads);
    for (int r = 0; r < ny; ++r)
        memcpy(field + (size_t)r * nx, g + (size_t)(r + oy) * Npx + ox, sizeof(c128) * nx);
    free(g);
}

/* ------------------------------------------------------------------------ */
/* Depth-map layer hologram (HologramGenerator, CPU path)                    */
/* ------------------------------------------------------------------------ */

/* Reference fftShift, DM/Hologram/src/HologramGenerator_CPU.cpp:332-344. */
static void ref_fftshift(int nx, int ny, const c128* in, c128* out) {
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j) {
            int ti = i - nx / 2; if (ti < 0) ti += nx;
            int tj = j - ny / 2; if (tj < 0) tj += ny;
            out[ti + (size_t)tj * nx] = in[i + (size_t)j * nx];
        }
}

/* fftwShift (HologramGenerator_CPU.cpp:285-322): shift, 2-D DFT (type 1 =
 * FFTW_FORWARD, -1 = FFTW_BACKWARD), optional 1/(nx ny), shift. */
static void ref_fftwshift(c128* a, int nx, int ny, int type, int normalized, int nthreads) {
    c128* t = (c128*)malloc(sizeof(c128) * (size_t)nx * ny);
    ref_fftshift(nx, ny, a, t);
    orc_fft2d(t, nx, ny, type == 1 ? -1 : +1, normalized, nthreads);
    ref_fftshift(nx, ny, t, a);
    free(t);
}

/* Depth levels as GetDepthValues builds them (HologramGenerator.cpp:393-401):
 * repeated addition of dstep from near.  The reference stops when the sum
 * exceeds far, which for some (near, far, L) drops the last level through
 * rounding drift and then reads past the list; here the sequence is
 * continued to exactly L entries. num_of_depth == 1 follows :405-406. */
void orc_depth_levels(double near_m, double far_m, int L, double* out) {
    if (L > 1) {
        const double dstep = (far_m - near_m) / (L - 1);
        double v = near_m;
        for (int i = 0; i < L; ++i) { out[i] = v; v += dstep; }
    } else {
        out[0] = far_m - near_m;
    }
}

/*
 * Depth index per pixel (0 = in no layer), the two modes of the reference:
 *  quant_mode 0 (FLAG_CHANGE_DEPTH_QUANTIZATION = 0): idx = Q - d,
 *    Q = DEFAULT_DEPTH_QUANTIZATION (HologramGenerator_CPU.cpp:64-65)
 *  quant_mode 1: dmap = (1 - d/255)(far - near) + near (CPU.cpp:60-62) and
 *    idx = sum_l (l+1) [d1_l <= dmap < d2_l] (<= d2 for the last level),
 *    d1,2 = level_l -/+ dstep/2 (change_depth_quan_CPU, CPU.cpp:74-103).
 */
void orc_depth_index(int n, const unsigned char* dep8, int quant_mode, int Q,
                     double near_m, double far_m, int L, int* idx) {
    if (quant_mode == 0) {
        for (int i = 0; i < n; ++i) idx[i] = Q - (int)dep8[i];
        return;
    }
    double* lev = (double*)malloc(sizeof(double) * L);
    orc_depth_levels(near_m, far_m, L, lev);
    const double dstep = L > 1 ? (far_m - near_m) / (L - 1) : (far_m + near_m) / 2;
    for (int i = 0; i < n; ++i) idx[i] = 0;
    for (int dtr = 0; dtr < L; ++dtr) {
        const double d1 = lev[dtr] - dstep / 2.0, d2 = lev[dtr] + dstep / 2.0;
        for (int i = 0; i < n; ++i) {
            const double dsrc = (double)dep8[i] / 255.0;
            const double dmap = (1 - dsrc) * (far_m - near_m) + near_m;
            int t;
            if (dtr < L - 1) t = (dmap >= d1 ? 1 : 0) * (dmap < d2 ? 1 : 0);
            else t = (dmap >= d1 ? 1 : 0) * (dmap <= d2 ? 1 : 0);
            idx[i] += t * (dtr + 1);
        }
    }
    free(lev);
}

/*
 * Symmetrisation encode of a centred-order spectrum U (Ny x Nx, overwritten),
 * Encoding_Symmetrization(ivec2(0,1)) of HologramGenerator.cpp:481-522 and
 * encoding_CPU / get_shift_phase_value of HologramGenerator_CPU.cpp:231-270,
 * 367-408:
 *   c = floor(Ny/2), keep 1-based spectrum rows c - floor(c/2) .. that + c - 1
 *   (all columns; :494-496, then -1 to 0-based :511-514), zero the rest;
 *   inverse fftwShift (FFTW_BACKWARD) normalised by 1/(Nx Ny);
 *   fringe[r][c] = Re(v[r + oy][c + ox] exp(i 2 pi yy / (4 py))),
 *   yy = Sy/2 - py r - py (CPU.cpp:376-390), Sy = h py.
 * The scene window (h x w) sits at offset ((Nx - w)/2, (Ny - h)/2) of the
 * grid; Nx = w, Ny = h is the reference's unpadded case.
 */
void orc_dm_encode(c128* U, int Nx, int Ny, int w, int h, double py, double* fringe,
                   int nthreads) {
    const int ox = (Nx - w) / 2, oy = (Ny - h) / 2;
    const int cropy = Ny / 2;
    const int cropy1 = cropy - cropy / 2 - 1;
    const int cropy2 = cropy1 + cropy - 1;
    for (int y = 0; y < Ny; ++y)
        if (y < cropy1 || y > cropy2) memset(U + (size_t)y * Nx, 0, sizeof(c128) * Nx);
    ref_fftwshift(U, Nx, Ny, -1, 1, nthreads);
    const double Sys = py * h;
    for (int r = 0; r < h; ++r) {
        const double yy = (Sys / 2.0) - py * r - py;
        const c128 sp = cexpi(2 * M_PI * (yy / (4 * py)));
        for (int c = 0; c < w; ++c)
            fringe[(size_t)r * w + c] = cmul(U[(size_t)(r + oy) * Nx + (c + ox)], sp).re;
    }
}

/*
 * Depth-map hologram of one colour, FP64, the CPU path of
 * DM/Hologram/src/HologramGenerator_CPU.cpp and HologramGenerator.cpp with
 * the ViewingWindow transform; depth index per orc_depth_index.
 *
 * Inputs: u8 intensity img and depth d, h x w (row-major).  The scene is
 * embedded (centred, offset ((Nx-w)/2, (Ny-h)/2)) in an Ny x Nx grid; Nx = w,
 * Ny = h reproduces the reference exactly (it does not pad).
 *   img = u8/255, alpha = [u8 > 0]                        (CPU.cpp:57-61)
 *   levels d_l (orc_depth_levels), z_l = -f d_l/(d_l - f) (HologramGenerator.cpp:432);
 *   field_lens <= 0 selects no field lens, z_l = d_l
 *   for layer l = 1..L: u_l = img*alpha*[idx == l] * rho * exp(i k z_l)
 *                                                         (CPU.cpp:139-156)
 *     U += H_l * fftwShift(u_l, forward)                 (CPU.cpp:159-160)
 *     H_l[i] = exp(i k (-z_l) sqrt(1 - (lambda fx)^2 - (lambda fy)^2)),
 *     fx = -1/(2px) + x/Sx,  fy = 1/(2py) - 1/Sy - y/Sy   (CPU.cpp:186-216)
 *   rho = (rho_re, rho_im): the same for every layer (get_rand_phase_value,
 *     HologramGenerator.cpp:457-473; (1, 0) when RANDOM_PHASE = 0).
 * Encode (Encoding_Symmetrization(0,1), HologramGenerator.cpp:481-522 and
 * encoding_CPU, CPU.cpp:231-270): keep spectrum rows [cropy1, cropy2]
 * (all columns), inverse fftwShift normalised by Nx Ny, crop the scene
 * window, multiply exp(i 2 pi yy/(4 py)), yy = Sy/2 - py r - py with r the
 * scene row and Sy = h py (CPU.cpp:376-390), keep Re.
 * Outputs: spectrum (Ny x Nx, centred order as U_complex_; may be NULL),
 * fringe (h x w, may be NULL) and field (h x w complex: the normalised
 * inverse transform of the whole spectrum in the scene window, i.e. the
 * hologram-plane field before encoding; may be NULL).  The CPU path's skipping of empty layers
 * (CPU.cpp:145) changes nothing numerically and is kept.
 */
void orc_depthmap(int w, int h, const unsigned char* img8, const unsigned char* dep8,
                  int Nx, int Ny, double px, double py, double lambda,
                  double near_m, double far_m, int L, int quant_mode, int Q,
                  double field_lens, double rho_re, double rho_im, int layer_lo,
                  int layer_hi, c128* spectrum, double* fringe, c128* field, int nthreads) {
    const size_t N = (size_t)Nx * Ny;
    int* didx = (int*)malloc(sizeof(int) * (size_t)w * h);
    orc_depth_index(w * h, dep8, quant_mode, Q, near_m, far_m, L, didx);
    const int ox = (Nx - w) / 2, oy = (Ny - h) / 2;
    const double k = TWO_PI / lambda;
    const double Sx = px * Nx, Sy = py * Ny;
    double* lev = (double*)malloc(sizeof(double) * L);
    orc_depth_levels(near_m, far_m, L, lev);
    c128* U = (c128*)calloc(N, sizeof(c128));
    c128* u = (c128*)malloc(sizeof(c128) * N);
    const c128 rho = {rho_re, rho_im};
    if (layer_lo < 1) layer_lo = 1;
    if (layer_hi > L || layer_hi < 1) layer_hi = L;
    for (int dtr = layer_lo; dtr <= layer_hi; ++dtr) {
        const double dl = lev[dtr - 1];
        const double zl = field_lens > 0.0 ? -field_lens * dl / (dl - field_lens) : dl;
        memset(u, 0, sizeof(c128) * N);
        double sum = 0.0;
        for (int r = 0; r < h; ++r)
            for (int c = 0; c < w; ++c) {
                const int i = r * w + c;
                const double v = img8[i] / 255.0 * (img8[i] > 0 ? 1.0 : 0.0) *
                                 (didx[i] == dtr ? 1.0 : 0.0);
                u[(size_t)(r + oy) * Nx + (c + ox)].re = v;
                sum += v;
            }
        if (!(sum > 0.0)) continue;
        double tc = k * zl / TWO_PI;
        tc -= floor(tc);
        const c128 carrier = cexpi(TWO_PI * tc);
        const c128 m = cmul(rho, carrier);
        for (size_t i = 0; i < N; ++i)
            if (u[i].re != 0.0) u[i] = cmul(u[i], m);
        ref_fftwshift(u, Nx, Ny, 1, 0, nthreads);
        const double dist = -zl;
        for (int y = 0; y < Ny; ++y) {
            const double fyy = (1.0 / (2.0 * py)) - (1.0 / Sy) - (1.0 / Sy) * y;
            for (int x = 0; x < Nx; ++x) {
                const double fxx = (-1.0 / (2.0 * px)) + (1.0 / Sx) * x;
                const double sq = sqrt(1 - (lambda * fxx) * (lambda * fxx) - (lambda * fyy) * (lambda * fyy));
                double t = dist * sq / lambda;  /* k dist sq in turns */
                t -= floor(t);
                const c128 kern = cexpi(TWO_PI * t);
                c128* e = &U[(size_t)y * Nx + x];
                const c128 p = cmul(kern, u[(size_t)y * Nx + x]);
                e->re += p.re; e->im += p.im;
            }
        }
    }
    if (spectrum) memcpy(spectrum, U, sizeof(c128) * N);
    if (field) {
        /* complex hologram field without the encoding: the normalised inverse
         * fftwShift of the whole spectrum, cropped to the scene window */
        c128* V = (c128*)malloc(sizeof(c128) * N);
        memcpy(V, U, sizeof(c128) * N);
        ref_fftwshift(V, Nx, Ny, -1, 1, nthreads);
        for (int r = 0; r < h; ++r)
            for (int c = 0; c < w; ++c) field[(size_t)r * w + c] = V[(size_t)(r + oy) * Nx + (c + ox)];
        free(V);
    }
    if (fringe) orc_dm_encode(U, Nx, Ny, w, h, py, fringe, nthreads);
    free(u);
    free(U);
    free(lev);
    free(didx);
}

/* ------------------------------------------------------------------------ */
/* Depth-map result encoding and eye-model reconstruction                    */
/* ------------------------------------------------------------------------ */

/*
 * 8-bit scaling of Write_Result_image (DM/Hologram/src/HologramGenerator.cpp:
 * 539-556): min/max with the reference's if / else-if update (equivalent to
 * plain min/max, since min <= max throughout), then
 * data[x] = (uint)((v - min)/(max - min) * 255) stored in an unsigned char.
 * A constant input (division 0/0 in the reference) gives 0 here.
 */
void orc_scale_u8(const double* src, unsigned char* dst, int64_t n) {
    double mn = src[0], mx = src[0];
    for (int64_t i = 0; i < n; ++i) {
        if (mn > src[i]) mn = src[i];
        else if (mx < src[i]) mx = src[i];
    }
    for (int64_t i = 0; i < n; ++i)
        dst[i] = mx > mn ? (unsigned char)(unsigned)((src[i] - mn) / (mx - mn) * 255) : 0;
}

/* HologramGenerator::circshift (HologramGenerator_CPU.cpp:663-680). */
static void ref_circshift(const c128* in, c128* out, int shift_x, int shift_y, int nx, int ny) {
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            int ti = (i + shift_x) % nx;
            if (ti < 0) ti += nx;
            int tj = (j + shift_y) % ny;
            if (tj < 0) tj += ny;
            out[ti + (size_t)tj * nx] = in[i + (size_t)j * nx];
        }
}

/*
 * Eye-model reconstruction, FP64, as ReconstructImage with
 * test_pixel_number_scale 1 and one focus distance
 * (HologramGenerator_CPU.cpp:415-614):
 *   Test_Propagation_to_Eye_Pupil (:496-532): hh = fftwShift(fringe, +1),
 *     pe = lambda f / (n p), xe = -E/2 + pe x, ye = (E_y/2 - pe_y) - pe_y y,
 *     hh_c = hh exp(i pi/lambda/f (xe^2 + ye^2));
 *   Reconstruction (:536-614): circshift(hh_c, -round(cx/pe_x), round(cy/pe_y)),
 *     f_eye = L (f - d)/(L + f - d), f_eff = f_eye L/(f_eye - L),
 *     hh_e = hh_s exp(i pi/lambda/f_eff r^2) [sqrt(r^2) < |lambda f_eff/(2 max pe)|]
 *            [sqrt(r^2) < pupil/2], fftwShift(hh_e, +1),
 *     sim_final = |hh_e exp(i pi/lambda/L (xr^2 + yr^2))|.
 * out: nx x ny.  Sizes: powers of two (the oracle's FFT).
 */
void orc_dm_reconstruct(const double* fringe, int nx, int ny, double ppx, double ppy, double lambda,
                        double f_field, double eye_length, double pupil, double cx, double cy,
                        double focus, double* out, int nthreads) {
    const size_t n = (size_t)nx * ny;
    c128* hh = (c128*)malloc(sizeof(c128) * n);
    c128* hs = (c128*)malloc(sizeof(c128) * n);
    for (size_t k = 0; k < n; ++k) { hh[k].re = fringe[k]; hh[k].im = 0.0; }
    ref_fftwshift(hh, nx, ny, 1, 0, nthreads);
    const double pp_ex = lambda * f_field / (nx * ppx), pp_ey = lambda * f_field / (ny * ppy);
    const double E_x = pp_ex * nx, E_y = pp_ey * ny;
    for (size_t p = 0; p < n; ++p) {
        const double x = (double)(p % nx), y = (double)(p / nx);
        const double xe = (-E_x / 2.0) + (pp_ex * x);
        const double ye = (E_y / 2.0 - pp_ey) - (pp_ey * y);
        hh[p] = cmul(hh[p], cexpi(M_PI / lambda / f_field * (xe * xe + ye * ye)));
    }
    const int sx = (int)lround(cx / pp_ex), sy = (int)lround(cy / pp_ey);
    ref_circshift(hh, hs, -sx, sy, nx, ny);
    const double f_eye = eye_length * (f_field - focus) / (eye_length + (f_field - focus));
    const double f_eff = f_eye * eye_length / (f_eye - eye_length);
    const double amax = fabs(lambda * f_eff / (2.0 * (pp_ex > pp_ey ? pp_ex : pp_ey)));
    for (size_t p = 0; p < n; ++p) {
        const double x = (double)(p % nx), y = (double)(p / nx);
        const double xe = (-E_x / 2.0) + (pp_ex * x);
        const double ye = (E_y / 2.0 - pp_ey) - (pp_ey * y);
        const double r = sqrt(xe * xe + ye * ye);
        const int m = (r < amax) && (r < pupil / 2.0);
        c128 v = cmul(hs[p], cexpi(M_PI / lambda / f_eff * (xe * xe + ye * ye)));
        hs[p].re = m ? v.re : 0.0;
        hs[p].im = m ? v.im : 0.0;
    }
    ref_fftwshift(hs, nx, ny, 1, 0, nthreads);
    const double pr_x = lambda * eye_length / E_x, pr_y = lambda * eye_length / E_y;
    const double R_x = pr_x * nx, R_y = pr_y * ny;
    for (size_t p = 0; p < n; ++p) {
        const double x = (double)(p % nx), y = (double)(p / nx);
        const double xr = (-R_x / 2.0) + (pr_x * x);
        const double yr = (R_y / 2.0 - pr_y) - (pr_y * y);
        const c128 v = cmul(hs[p], cexpi(M_PI / lambda / eye_length * (xr * xr + yr * yr)));
        out[p] = sqrt(v.re * v.re + v.im * v.im);
    }
    free(hh);
    free(hs);
}
