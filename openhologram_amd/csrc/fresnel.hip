// fresnel.hip -- Fresnel point-cloud hologram on gfx950: a plain FP32 VALU
// kernel and the separable form as an MFMA complex contraction.
//
// Field (the direct Fresnel sum of CWO/cwoPLS.cpp:370-422, with the point
// amplitude, its initial phase and the plane carrier exp(i k z_n)):
//   U[i][j] = sum_n a_n exp(i (k z_n + phi_n)) exp(i k ((X_j - x_n)^2 + (Y_i - y_n)^2) / (2 z_n))
//   X_j = (j - nx/2) p,  Y_i = (i - ny/2) p                 (cwoPLS.cpp:400-401)
// The quadratic phase separates, exp(i k (dx^2 + dy^2)/(2z)) =
// exp(i k dx^2/(2z)) exp(i k dy^2/(2z)), so U = A B^T with
//   A[i][n] = w_n exp(i pi (Y_i - y_n)^2/(lambda z_n)),  B[j][n] = exp(i pi (X_j - x_n)^2/(lambda z_n)),
//   w_n = a_n exp(i (k z_n + phi_n)):
// a complex (pixels x points)(points x pixels) contraction.  The chirp phases
// reach ~1e4 turns and are reduced in FP64 before the FP32 sin/cos; the
// factors are rounded to BF16 or FP16, optionally split as hi + lo (the lo
// part carries the rounding error of hi), and contracted on MFMA with FP32
// accumulation (A B = Ah Bh + Ah Bl + Al Bh).
#include <algorithm>
#include <cmath>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "oph_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

namespace {

struct FrPoint {
    double x, y, c;   // position, c = 1/(2 lambda z) (turns per m^2)
    float wre, wim;   // a exp(i 2 pi frac(z/lambda + phi/2pi))
    float f0, a;      // frac(z/lambda + phi/2pi), amplitude
};

__device__ FrPoint make_frpoint(const double* x, const double* y, const double* z, const double* amp,
                                const double* phase, int64_t i, double lambda) {
    FrPoint q;
    q.x = x[i];
    q.y = y[i];
    q.c = 1.0 / (2.0 * lambda * z[i]);
    double t = z[i] / lambda + (phase ? phase[i] * 0.15915494309189535 : 0.0);
    t -= rint(t);
    const float a = amp ? (float)amp[i] : 1.f;
    q.f0 = (float)t;
    q.a = a;
    q.wre = a * cos_turns((float)t);
    q.wim = a * sin_turns((float)t);
    return q;
}

// a zero-weight point (padding up to whole contraction chunks) on plane c
__device__ FrPoint zero_frpoint(double c) {
    FrPoint q;
    q.x = q.y = 0.0;
    q.c = c;
    q.f0 = 0.f;
    q.a = 0.f;
    q.wre = q.wim = 0.f;
    return q;
}

// skip: when set and *skip != 0 the plane-factored form computes the field
// and this general preparation (and contraction) does nothing
__global__ void fresnel_prep(const double* x, const double* y, const double* z, const double* amp,
                             const double* phase, int64_t n, int64_t npad, double lambda, FrPoint* out,
                             const int* skip) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npad || (skip && *skip)) return;
    out[i] = i < n ? make_frpoint(x, y, z, amp, phase, i, lambda) : zero_frpoint(0.0);
}

// ---------------------------------------------------------------------------
// plain kernel: a 64x64 pixel tile per workgroup, 4x4 pixels per thread,
// points staged in LDS; per (point, thread) the phase at the block centre in
// FP64, per pixel the exact quadratic offset c (2 ex dx0 + ex^2 + 2 ey dy0 + ey^2)
// in FP32 (a few turns at most).  Each staged batch of PT points is summed
// in FP32 and the batch sums in FP64, so a pixel's rounding error does not
// grow with the point count (1e6 points in one sum: ~2e-5 relative in FP32).

constexpr int PT = 256;

struct ValuArgs {
    const FrPoint* pts;
    int64_t n, chunk;
    int nx, ny;
    double p;
    float2* partial;
};

__global__ __launch_bounds__(256) void fresnel_valu(ValuArgs a) {
    __shared__ double s_x[PT], s_y[PT], s_c[PT];
    __shared__ float s_f0[PT], s_a[PT];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 16 + tx;
    const int c0 = blockIdx.x * 64 + tx * 4, r0 = blockIdx.y * 64 + ty * 4;
    const double X0 = (c0 + 1.5 - a.nx / 2) * a.p, Y0 = (r0 + 1.5 - a.ny / 2) * a.p;
    float ex[4], ey[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        ex[k] = (float)((k - 1.5) * a.p);
        ey[k] = (float)((k - 1.5) * a.p);
    }
    double sre[4][4], sim[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) sre[r][c] = sim[r][c] = 0.0;
    const int64_t pb = (int64_t)blockIdx.z * a.chunk;
    const int64_t pe = min(a.n, pb + a.chunk);
    for (int64_t base = pb; base < pe; base += PT) {
        __syncthreads();
        const int64_t j = base + tid;
        if (j < pe) {
            const FrPoint q = a.pts[j];
            s_x[tid] = q.x;
            s_y[tid] = q.y;
            s_c[tid] = q.c;
            s_f0[tid] = q.f0;
            s_a[tid] = q.a;
        }
        __syncthreads();
        const int cnt = (int)min((int64_t)PT, pe - base);
        float re[4][4], im[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) re[r][c] = im[r][c] = 0.f;
        for (int k = 0; k < cnt; ++k) {
            const double dx0 = X0 - s_x[k], dy0 = Y0 - s_y[k];
            const double cc = s_c[k];
            double t = fma(cc, fma(dx0, dx0, dy0 * dy0), (double)s_f0[k]);
            const float f0 = (float)(t - rint(t));
            const float cf = (float)cc, am = s_a[k];
            const float dxf = (float)dx0, dyf = (float)dy0;
            float A[4], B[4];
#pragma unroll
            for (int c = 0; c < 4; ++c) A[c] = cf * fmaf(2.f * ex[c], dxf, ex[c] * ex[c]);
#pragma unroll
            for (int r = 0; r < 4; ++r) B[r] = cf * fmaf(2.f * ey[r], dyf, ey[r] * ey[r]);
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const float ph = f0 + (A[c] + B[r]);
                    re[r][c] = fmaf(am, cos_turns(ph), re[r][c]);
                    im[r][c] = fmaf(am, sin_turns(ph), im[r][c]);
                }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                sre[r][c] += (double)re[r][c];
                sim[r][c] += (double)im[r][c];
            }
    }
    float2* dst = a.partial + (size_t)blockIdx.z * a.nx * a.ny;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = r0 + r;
        if (row >= a.ny) continue;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int col = c0 + c;
            if (col < a.nx) dst[(size_t)row * a.nx + col] = make_float2((float)sre[r][c], (float)sim[r][c]);
        }
    }
}

__global__ void fresnel_sum_chunks(const float2* partial, int chunks, int64_t npix, float2* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    double re = 0.0, im = 0.0;
    for (int c = 0; c < chunks; ++c) {
        const float2 v = partial[(size_t)c * npix + i];
        re += v.x;
        im += v.y;
    }
    out[i] = make_float2((float)re, (float)im);
}

// ---------------------------------------------------------------------------
// MFMA contraction.  A workgroup of 8 waves owns a 256 x 128 output tile
// (256 rows of y, 128 columns of x); for each chunk of 32 points it writes
// the chirp factors A (256 rows) and B (128 columns) to LDS as 16-bit planes
// [re_hi, im_hi, re_lo, im_lo] laid out [row][point] (generated by a phase
// recurrence, below), then each wave contracts a 64 x 64 sub-tile with
// v_mfma_f32_32x32x16 (lane l supplies row l%32, points 8(l/32)..+7 of A and
// column l%32, the same points of B; accumulator register v of lane l is
// output row 8(v/4) + 4(l/32) + v%4, column l%32).  A chirp entry is
// generated once per tile and used by every output of its row or column, so
// the taller tile spends 3/4 of the generation per output of a square
// 128 x 128 one; the workgroup takes the whole CU (two waves per SIMD).

constexpr int MR = 256;         // output tile rows
constexpr int MC = 128;         // output tile columns
constexpr int KC = 32;          // padding granularity: the largest chunk
// points per chunk: 32 for the single-precision forms; 16 for the hi+lo
// forms, whose four operand planes in two buffers would not fit LDS at 32
template <bool SPLIT>
constexpr int KCH = SPLIT ? 16 : 32;
constexpr int MFMA_THREADS = 512;

template <typename T>
struct Mfma;

template <>
struct Mfma<__bf16> {
    typedef bf16x8 V;
    static __device__ __forceinline__ __bf16 cvt(float f) { return (__bf16)f; }
    // both parts in one packed conversion (round to nearest even)
    static __device__ __forceinline__ void cvt2(float re, float im, __bf16& rh, __bf16& ih) {
        const f32x2 v = {re, im};
        const bf16x2 h = __builtin_convertvector(v, bf16x2);
        rh = h.x;
        ih = h.y;
    }
    static __device__ __forceinline__ float back(__bf16 h) { return (float)h; }
    // the low parts (re - rh, im - ih, exact in FP32) in one packed conversion
    static __device__ __forceinline__ void lo2(float re, float im, __bf16 rh, __bf16 ih, float, __bf16& rl,
                                               __bf16& il) {
        cvt2(re - (float)rh, im - (float)ih, rl, il);
    }
    static __device__ __forceinline__ f32x16 mac(V a, V b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

template <>
struct Mfma<_Float16> {
    typedef f16x8 V;
    static __device__ __forceinline__ _Float16 cvt(float f) { return (_Float16)f; }
    static __device__ __forceinline__ void cvt2(float re, float im, _Float16& rh, _Float16& ih) {
        const f32x2 v = {re, im};
        const f16x2 h = __builtin_convertvector(v, f16x2);
        rh = h.x;
        ih = h.y;
    }
    static __device__ __forceinline__ float back(_Float16 h) { return (float)h; }
    // f - h rounded to FP16 (exact in FP32 before the rounding) as
    // f + h m1 with m1 = -1 from a register (a literal -1 would fold the FMA
    // into a subtraction): one mixed-precision FMA, FP16 operand widened
    static __device__ __forceinline__ _Float16 lo(float f, _Float16 h, float m1) {
        return (_Float16)fmaf((float)h, m1, f);
    }
    static __device__ __forceinline__ void lo2(float re, float im, _Float16 rh, _Float16 ih, float m1, _Float16& rl,
                                               _Float16& il) {
        rl = lo(re, rh, m1);
        il = lo(im, ih, m1);
    }
    static __device__ __forceinline__ f32x16 mac(V a, V b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};

// two 16-bit values as one word, the first in the low half
template <typename T>
__device__ __forceinline__ unsigned pack2(T lo, T hi) {
    return (unsigned)__builtin_bit_cast(unsigned short, lo) | ((unsigned)__builtin_bit_cast(unsigned short, hi) << 16);
}

template <typename V>
__device__ __forceinline__ V negate(V v) {
    u32x4 u = __builtin_bit_cast(u32x4, v);
    u ^= 0x80008000u;
    return __builtin_bit_cast(V, u);
}

struct MfmaArgs {
    const FrPoint* pts;
    int64_t npad;     // multiple of KC
    int nx, ny;
    double p;
    float2* out;
    const int* skip;  // when set and nonzero, the plane-factored form has the field
};

template <typename T, bool SPLIT>
struct FrLds {
    static constexpr int NP = SPLIT ? 4 : 2;
    static constexpr int KS = KCH<SPLIT> + 8;   // row stride (elements): 16-byte aligned rows
    T A[NP][MR][KS];
    T B[NP][MC][KS];
};

// Chirp factors by a phase recurrence.  A run is one point's entries on
// rows (A, with the point weight) or columns (B) gph + G m, G = 32.  Along a
// run the phase c (Y - y)^2 has constant second difference
// delta = 2 c (G p)^2, so with R = exp(i 2 pi (phi_{m+1} - phi_m)),
// A_{m+1} = A_m R_m and R_{m+1} = R_m exp(i 2 pi delta): two complex
// multiplies per entry; the anchor phi_0, the first step and delta are
// reduced in FP64 once per run (phases of ~1e4 turns).  In the
// plane-factored form (below) the factor is geometric: one multiply.
// FULL: every row and column of the tile is inside the image (no per-entry
// bounds select).
//
// Two neighbouring points per thread: a thread runs points k and k + 1 over
// the same rows, so each operand plane takes one 32-bit LDS store per row
// (both points' 16-bit parts) instead of two 16-bit stores.  The generation
// is bound by the LDS store instructions, not by its arithmetic; the extra
// FP64 set-up of the second run is the cheaper side.
constexpr int GPH2 = 32;   // run stride with point pairs

template <typename T, bool SPLIT, bool FULL, int AB, bool PLANE, typename Args>
__device__ __forceinline__ void fr_run_pair(FrLds<T, SPLIT>& L, const double* p_x, const double* p_y,
                                            const double* p_c, const float* p_wr, const float* p_wi, int k, int gph,
                                            int i0, int j0, const Args& a) {
    constexpr int M = (AB == 0 ? MR : MC) / GPH2;
    constexpr int KS = FrLds<T, SPLIT>::KS;
    constexpr int PS = (AB == 0 ? MR : MC) * KS / 2;   // plane stride in 32-bit words
    const double sp = (double)GPH2 * a.p;
    const double P0 = AB == 0 ? (i0 + gph - a.ny / 2) * a.p : (j0 + gph - a.nx / 2) * a.p;
    float2 v[2], R[2], E[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const double c = p_c[k + h];
        const double u = AB == 0 ? p_y[k + h] : p_x[k + h];
        double t0, t1;
        if constexpr (PLANE) {
            // the plane's factor: phase c (u^2 - 2 P u), step -2 c u G p
            t0 = c * u * (u - 2.0 * P0);
            t1 = -2.0 * c * u * sp;
        } else {
            // the full chirp: phase c d^2, first step c (2 d G p + (G p)^2),
            // second difference 2 c (G p)^2
            const double d0 = P0 - u;
            t0 = c * d0 * d0;
            t1 = c * (2.0 * d0 * sp + sp * sp);
            double td = 2.0 * c * sp * sp;
            td -= rint(td);
            E[h] = make_float2(cos_turns((float)td), sin_turns((float)td));
        }
        t0 -= rint(t0);
        t1 -= rint(t1);
        v[h] = make_float2(cos_turns((float)t0), sin_turns((float)t0));
        if constexpr (AB == 0) {
            const float wr_ = p_wr[k + h], wi_ = p_wi[k + h];
            v[h] = make_float2(wr_ * v[h].x - wi_ * v[h].y, wr_ * v[h].y + wi_ * v[h].x);
        }
        R[h] = make_float2(cos_turns((float)t1), sin_turns((float)t1));
    }
    const int lim = AB == 0 ? a.ny - i0 : a.nx - j0;   // rows / columns inside the image
    unsigned* p0 = reinterpret_cast<unsigned*>(AB == 0 ? &L.A[0][gph][k] : &L.B[0][gph][k]);
    float m1 = -1.f;
    if constexpr (SPLIT) asm volatile("v_mov_b32 %0, -1.0" : "=v"(m1));
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const bool ok = FULL || gph + GPH2 * m < lim;
        float re[2], im[2];
        T rh[2], ih[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            re[h] = ok ? v[h].x : 0.f;
            im[h] = ok ? v[h].y : 0.f;
        }
        Mfma<T>::cvt2(re[0], re[1], rh[0], rh[1]);
        Mfma<T>::cvt2(im[0], im[1], ih[0], ih[1]);
        unsigned* q = p0 + m * GPH2 * KS / 2;
        q[0] = pack2(rh[0], rh[1]);
        q[PS] = pack2(ih[0], ih[1]);
        if constexpr (SPLIT) {
            T rl[2], il[2];
            Mfma<T>::lo2(re[0], re[1], rh[0], rh[1], m1, rl[0], rl[1]);
            Mfma<T>::lo2(im[0], im[1], ih[0], ih[1], m1, il[0], il[1]);
            q[2 * PS] = pack2(rl[0], rl[1]);
            q[3 * PS] = pack2(il[0], il[1]);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            v[h] = make_float2(fmaf(v[h].x, R[h].x, -v[h].y * R[h].y), fmaf(v[h].x, R[h].y, v[h].y * R[h].x));
            if constexpr (!PLANE)
                R[h] = make_float2(fmaf(R[h].x, E[h].x, -R[h].y * E[h].y), fmaf(R[h].x, E[h].y, R[h].y * E[h].x));
        }
    }
}

// One thread per (point pair, phase) -- or, with 16-point chunks, per
// (point pair, phase, factor): the first four waves take the A runs and the
// last four the B runs, one of each per SIMD.
template <typename T, bool SPLIT, bool FULL, bool PLANE, typename Args>
__device__ __forceinline__ void fr_generate_pairs(FrLds<T, SPLIT>& L, const double* p_x, const double* p_y,
                                                  const double* p_c, const float* p_wr, const float* p_wi, int tid,
                                                  int i0, int j0, const Args& a) {
    constexpr int NPAIR = KCH<SPLIT> / 2;
    if constexpr (!SPLIT) {
        static_assert(NPAIR * GPH2 == MFMA_THREADS, "one thread per (point pair, phase)");
        const int k = 2 * (tid % NPAIR), gph = tid / NPAIR;
        fr_run_pair<T, SPLIT, FULL, 0, PLANE>(L, p_x, p_y, p_c, p_wr, p_wi, k, gph, i0, j0, a);
        fr_run_pair<T, SPLIT, FULL, 1, PLANE>(L, p_x, p_y, p_c, p_wr, p_wi, k, gph, i0, j0, a);
    } else {
        static_assert(2 * NPAIR * GPH2 == MFMA_THREADS, "one thread per (point pair, phase, factor)");
        const int h = tid % (MFMA_THREADS / 2);
        const int k = 2 * (h % NPAIR), gph = h / NPAIR;
        if (tid < MFMA_THREADS / 2)
            fr_run_pair<T, SPLIT, FULL, 0, PLANE>(L, p_x, p_y, p_c, p_wr, p_wi, k, gph, i0, j0, a);
        else
            fr_run_pair<T, SPLIT, FULL, 1, PLANE>(L, p_x, p_y, p_c, p_wr, p_wi, k, gph, i0, j0, a);
    }
}

// Contraction of one chunk from L by wave (wr, wc): its 64 x 64 sub-tile.
template <typename T, bool SPLIT>
__device__ __forceinline__ void fr_contract(const FrLds<T, SPLIT>& L, int wr, int wc, int lane, f32x16 (&cr)[2][2],
                                            f32x16 (&ci)[2][2]) {
    typedef typename Mfma<T>::V V;
    constexpr int NP = FrLds<T, SPLIT>::NP;
#pragma unroll
    for (int s = 0; s < KCH<SPLIT> / 16; ++s) {
        const int kb = 16 * s + 8 * (lane >> 5);
        V A[2][NP], B[2][NP];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            const int row = wr * 64 + 32 * x + (lane & 31);
            const int col = wc * 64 + 32 * x + (lane & 31);
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                A[x][q] = *reinterpret_cast<const V*>(&L.A[q][row][kb]);
                B[x][q] = *reinterpret_cast<const V*>(&L.B[q][col][kb]);
            }
        }
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const V nbi = negate(B[y][1]);
            const V nbil = SPLIT ? negate(B[y][NP - 1]) : nbi;
#pragma unroll
            for (int x = 0; x < 2; ++x) {
                // re: Ar Br - Ai Bi ; im: Ar Bi + Ai Br  (hi x hi)
                cr[x][y] = Mfma<T>::mac(A[x][0], B[y][0], cr[x][y]);
                cr[x][y] = Mfma<T>::mac(A[x][1], nbi, cr[x][y]);
                ci[x][y] = Mfma<T>::mac(A[x][0], B[y][1], ci[x][y]);
                ci[x][y] = Mfma<T>::mac(A[x][1], B[y][0], ci[x][y]);
                if constexpr (SPLIT) {
                    // hi x lo and lo x hi
                    cr[x][y] = Mfma<T>::mac(A[x][0], B[y][2], cr[x][y]);
                    cr[x][y] = Mfma<T>::mac(A[x][1], nbil, cr[x][y]);
                    cr[x][y] = Mfma<T>::mac(A[x][2], B[y][0], cr[x][y]);
                    cr[x][y] = Mfma<T>::mac(A[x][3], nbi, cr[x][y]);
                    ci[x][y] = Mfma<T>::mac(A[x][0], B[y][3], ci[x][y]);
                    ci[x][y] = Mfma<T>::mac(A[x][1], B[y][2], ci[x][y]);
                    ci[x][y] = Mfma<T>::mac(A[x][2], B[y][1], ci[x][y]);
                    ci[x][y] = Mfma<T>::mac(A[x][3], B[y][0], ci[x][y]);
                }
            }
        }
    }
}

// The factors of chunk c + 1 are generated into one LDS buffer while chunk c
// is contracted from the other: one barrier per chunk.  The points reach
// LDS two chunks ahead (double-buffered too) and are loaded into registers
// three chunks ahead.
template <typename T, bool SPLIT>
__global__ __launch_bounds__(MFMA_THREADS) void fresnel_mfma(MfmaArgs a) {
    typedef FrLds<T, SPLIT> Lds;
    constexpr int K = KCH<SPLIT>;
    extern __shared__ __attribute__((aligned(16))) unsigned char fr_dyn[];
    Lds* L = reinterpret_cast<Lds*>(fr_dyn);   // two buffers
    __shared__ double p_x[2][K], p_y[2][K], p_c[2][K];
    __shared__ float p_wr[2][K], p_wi[2][K];

    if (a.skip && *a.skip) return;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wr = wave >> 1, wc = wave & 1;
    // XCD-aware order of the tiles: consecutive logical tiles on one XCD
    const int ntx = (a.nx + MC - 1) / MC;
    const int g = xcd_remap(blockIdx.x, gridDim.x);
    const int i0 = (g / ntx) * MR, j0 = (g % ntx) * MC;
    const bool full = i0 + MR <= a.ny && j0 + MC <= a.nx;
    const int64_t nch = a.npad / K;

    f32x16 cr[2][2], ci[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            cr[x][y] = (f32x16)0.f;
            ci[x][y] = (f32x16)0.f;
        }
    auto stage_points = [&](int slot, const FrPoint& q) {
        p_x[slot][tid] = q.x;
        p_y[slot][tid] = q.y;
        p_c[slot][tid] = q.c;
        p_wr[slot][tid] = q.wre;
        p_wi[slot][tid] = q.wim;
    };
    auto generate = [&](int64_t c) {
        const int slot = (int)(c & 1);
        if (full)
            fr_generate_pairs<T, SPLIT, true, false>(L[slot], p_x[slot], p_y[slot], p_c[slot], p_wr[slot], p_wi[slot],
                                                     tid, i0, j0, a);
        else
            fr_generate_pairs<T, SPLIT, false, false>(L[slot], p_x[slot], p_y[slot], p_c[slot], p_wr[slot],
                                                      p_wi[slot], tid, i0, j0, a);
    };
    FrPoint nq{};
    if (tid < K) {
        if (nch > 0) stage_points(0, a.pts[tid]);
        if (nch > 1) stage_points(1, a.pts[K + tid]);
        if (nch > 2) nq = a.pts[2 * K + tid];
    }
    __syncthreads();
    if (nch > 0) generate(0);
    __syncthreads();
    for (int64_t c = 0; c < nch; ++c) {
        if (c + 1 < nch) generate(c + 1);
        fr_contract<T, SPLIT>(L[c & 1], wr, wc, lane, cr, ci);
        // slot c & 1 held chunk c's points, read by generate(c) one step ago
        if (tid < K && c + 2 < nch) {
            stage_points((int)(c & 1), nq);
            if (c + 3 < nch) nq = a.pts[(c + 3) * K + tid];
        }
        __syncthreads();
    }
    // epilogue
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int row = i0 + wr * 64 + 32 * x + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const int col = j0 + wc * 64 + 32 * y + (lane & 31);
                if (row < a.ny && col < a.nx) a.out[(size_t)row * a.nx + col] = make_float2(cr[x][y][v], ci[x][y][v]);
            }
}

template <typename T, bool SPLIT>
int launch_fresnel_mfma(oph_ctx* ctx, const MfmaArgs& m, unsigned tiles) {
    const size_t shm = 2 * sizeof(FrLds<T, SPLIT>);
    static const bool attr = [&] {
        hipFuncSetAttribute(reinterpret_cast<const void*>(&fresnel_mfma<T, SPLIT>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        hipGetLastError();
        return true;
    }();
    (void)attr;
    OPH_LAUNCH(ctx, "fresnel_mfma", (fresnel_mfma<T, SPLIT>), dim3(tiles), dim3(MFMA_THREADS), shm, m);
    return OPH_OK;
}

// ---------------------------------------------------------------------------
// Plane-factored contraction.  Points on one plane z_p share
// c_p = 1/(2 lambda z_p), and
//   A[i][n] = D_p(Y_i) w_n exp(i 2 pi c_p (y_n^2 - 2 Y_i y_n)),  D_p(Y) = exp(i 2 pi c_p Y^2),
// B likewise with X_j and x_n (no weight).  D_p is common to the plane's
// points, so the plane's part of the field is
//   U_p[i][j] = D_p(Y_i) D_p(X_j) sum_{n on p} g_n[i] h_n[j],
// and g, h are geometric along a run of rows (ratio exp(-i 2 pi 2 c_p y_n G p)):
// one complex multiply per entry instead of two.  The points are grouped by
// plane (a stable sort by z, so a plane's points keep their input order; each
// plane padded to whole 32-point chunks with zero-weight points), the
// contraction runs plane by plane, and at each plane's last chunk the
// accumulators are multiplied by D_p(Y_i) D_p(X_j) and added to the output.
// Partial sums per plane also keep the FP32 accumulation error of a long
// sum to the plane's share.  Used when the points lie on at most PLANE_MAX
// distinct planes with at least PLANE_MIN_AVG points per plane on average
// (the device decides after grouping, and the general kernel runs otherwise).
constexpr int PLANE_MAX = 4096;
constexpr int PLANE_MIN_AVG = 256;
constexpr int64_t PLANE_D_MAX = (int64_t)1 << 25;   // entries of the plane-chirp tables

__global__ void fr_keys(const double* z, int64_t n, unsigned long long* key, unsigned* idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = (unsigned long long)__double_as_longlong(z[i]);
    idx[i] = (unsigned)i;
}

__global__ void fr_flags(const unsigned long long* key, int64_t n, unsigned* flag) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || key[i] != key[i - 1]) ? 1u : 0u;
}

// pid[i] = 1 + plane of sorted point i; start of each plane, its c, the count
__global__ void fr_starts(const unsigned long long* key, const unsigned* flag, const unsigned* pid, int64_t n,
                          double lambda, unsigned* start, double* pc, int* info) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flag[i]) {
        const unsigned p = pid[i] - 1;
        if (p < (unsigned)PLANE_MAX) {
            start[p] = (unsigned)i;
            pc[p] = 1.0 / (2.0 * lambda * __longlong_as_double((long long)key[i]));
        }
    }
    if (i == n - 1) info[1] = (int)min(pid[i], (unsigned)PLANE_MAX + 1);
}

// One workgroup: decides the form (info[0]) and pads each plane to whole
// chunks (poff: exclusive prefix of the padded counts).
__global__ __launch_bounds__(1024) void fr_plan(int64_t n, const unsigned* start, unsigned* poff, int* info,
                                                int64_t cap_chunks, int64_t dlen) {
    __shared__ unsigned part[1024];
    const int np = info[1];
    const int t = threadIdx.x;
    if (!(np >= 1 && np <= PLANE_MAX && (int64_t)np * PLANE_MIN_AVG <= n && (int64_t)np * dlen <= PLANE_D_MAX)) {
        if (t == 0) info[0] = 0;
        return;
    }
    unsigned pad[4], sum = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int q = 4 * t + j;
        const unsigned cnt = q < np ? (q + 1 < np ? start[q + 1] : (unsigned)n) - start[q] : 0u;
        pad[j] = (cnt + KC - 1) / KC * KC;
        sum += pad[j];
    }
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const unsigned v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    const unsigned total = part[1023];
    if ((int64_t)(total / KC) > cap_chunks) {
        if (t == 0) info[0] = 0;
        return;
    }
    unsigned run = part[t] - sum;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int q = 4 * t + j;
        if (q < np) poff[q] = run;
        run += pad[j];
    }
    if (t == 0) {
        poff[np] = total;
        info[2] = (int)(total / KC);
        info[0] = 1;
    }
}

// grouped points: sorted point i to poff[p] + (i - start[p]); the tail of
// each plane's last chunk gets zero-weight points
__global__ void fr_fill(const double* x, const double* y, const double* z, const double* amp, const double* phase,
                        int64_t n, double lambda, const unsigned* order, const unsigned* pid, const unsigned* start,
                        const unsigned* poff, const double* pc, const int* info, FrPoint* pts) {
    if (!info[0]) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int np = info[1];
    if (i < n) {
        const unsigned p = pid[i] - 1;
        pts[poff[p] + (unsigned)(i - start[p])] = make_frpoint(x, y, z, amp, phase, order[i], lambda);
    }
    if (i < (int64_t)np * KC) {
        const int p = (int)(i / KC), r = (int)(i % KC);
        const unsigned cnt = (p + 1 < np ? start[p + 1] : (unsigned)n) - start[p];
        if (poff[p] + cnt + r < poff[p + 1]) pts[poff[p] + cnt + r] = zero_frpoint(pc[p]);
    }
}

// plane chirps D_p(Y_i) = exp(i 2 pi c_p Y_i^2) for the rows (dr[p ny + i])
// and D_p(X_j) for the columns (dc[p nx + j]), phases reduced in FP64
__global__ void fr_chirps(const double* pc, const int* info, int nx, int ny, double pitch, float2* dr, float2* dc) {
    if (!info[0]) return;
    const int np = info[1];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nr = (int64_t)np * ny, ncl = (int64_t)np * nx;
    if (i >= nr + ncl) return;
    const bool row = i < nr;
    const int64_t e = row ? i : i - nr;
    const int len = row ? ny : nx;
    const int p = (int)(e / len), k = (int)(e % len);
    const double X = (k - len / 2) * pitch;
    double t = pc[p] * X * X;
    t -= rint(t);
    const float2 v = make_float2(cos_turns((float)t), sin_turns((float)t));
    if (row)
        dr[e] = v;
    else
        dc[e] = v;
}

struct PlaneMfmaArgs {
    const FrPoint* pts;   // grouped by plane, padded
    const unsigned* poff; // padded offset of each plane's first point (np + 1)
    const double* pc;     // c of each plane
    const int* info;      // [0] this form on, [1] planes, [2] 32-point chunks
    const float2* dr;     // plane chirps of the rows, [p ny + i]
    const float2* dc;     // and of the columns, [p nx + j]
    int nx, ny;
    double p;
    float2* out;
};

// The end of a plane: out (+)= D(Y) D(X) acc over the wave's 64 x 64 sub-tile
// (stored on the tile's first plane, added on the others), then acc = 0.
__device__ __forceinline__ void fr_flush(f32x16 (&cr)[2][2], f32x16 (&ci)[2][2], int pl, int wr, int wc, int lane,
                                         int i0, int j0, const PlaneMfmaArgs& a) {
    const bool first = pl == 0;
    // the lane index through a zero the compiler cannot see through: the
    // flush's addresses are the same at every plane and, hoisted out of the
    // chunk loop, would hold ~70 registers across it
    int zero;
    asm volatile("v_mov_b32 %0, 0" : "=v"(zero));
    lane += zero;
    float2 dc[2];
    int col[2];
#pragma unroll
    for (int y = 0; y < 2; ++y) {
        col[y] = j0 + wc * 64 + 32 * y + (lane & 31);
        dc[y] = col[y] < a.nx ? a.dc[(size_t)pl * a.nx + col[y]] : make_float2(0.f, 0.f);
    }
    // four rows at a time: their chirps and previous outputs are all
    // requested before any is used (addresses clamped into the image, so no
    // branch guards a load; stores outside the image are skipped)
    const int cc0 = min(col[0], a.nx - 1), cc1 = min(col[1], a.nx - 1);
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int vg = 0; vg < 4; ++vg) {
            float2 dr[4], pv[4][2];
            int row[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                row[q] = i0 + wr * 64 + 32 * x + 8 * vg + 4 * (lane >> 5) + q;
                const int rc = min(row[q], a.ny - 1);
                dr[q] = a.dr[(size_t)pl * a.ny + rc];
                // (loaded on the first plane too, then not used: no branch)
                pv[q][0] = a.out[(size_t)rc * a.nx + cc0];
                pv[q][1] = a.out[(size_t)rc * a.nx + cc1];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                asm volatile("" : "+v"(dr[q].x), "+v"(dr[q].y), "+v"(pv[q][0].x), "+v"(pv[q][0].y), "+v"(pv[q][1].x),
                             "+v"(pv[q][1].y));
                if (first) pv[q][0] = pv[q][1] = make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int v = 4 * vg + q;
#pragma unroll
                for (int y = 0; y < 2; ++y) {
                    const float2 d = make_float2(dr[q].x * dc[y].x - dr[q].y * dc[y].y,
                                                 dr[q].x * dc[y].y + dr[q].y * dc[y].x);
                    const float re = cr[x][y][v], im = ci[x][y][v];
                    const float2 val =
                        make_float2(pv[q][y].x + (re * d.x - im * d.y), pv[q][y].y + (re * d.y + im * d.x));
                    if (row[q] < a.ny && col[y] < a.nx) a.out[(size_t)row[q] * a.nx + col[y]] = val;
                }
            }
        }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            cr[x][y] = (f32x16)0.f;
            ci[x][y] = (f32x16)0.f;
        }
}

template <typename T, bool SPLIT>
__global__ __launch_bounds__(MFMA_THREADS) void fresnel_mfma_planes(PlaneMfmaArgs a) {
    typedef FrLds<T, SPLIT> Lds;
    constexpr int K = KCH<SPLIT>;
    if (!a.info[0]) return;
    extern __shared__ __attribute__((aligned(16))) unsigned char fr_dyn[];
    Lds* L = reinterpret_cast<Lds*>(fr_dyn);   // two buffers
    __shared__ double p_x[2][K], p_y[2][K], p_c[2][K];
    __shared__ float p_wr[2][K], p_wi[2][K];

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wr = wave >> 1, wc = wave & 1;
    const int ntx = (a.nx + MC - 1) / MC;
    const int g = xcd_remap(blockIdx.x, gridDim.x);
    const int i0 = (g / ntx) * MR, j0 = (g % ntx) * MC;
    const bool full = i0 + MR <= a.ny && j0 + MC <= a.nx;
    const int64_t nch = (int64_t)a.info[2] * (KC / K);
    // the current plane and its last K-point chunk
    int pl = 0;
    int64_t pend = (int64_t)(a.poff[1] / K) - 1;

    f32x16 cr[2][2], ci[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            cr[x][y] = (f32x16)0.f;
            ci[x][y] = (f32x16)0.f;
        }
    auto stage_points = [&](int slot, const FrPoint& q) {
        p_x[slot][tid] = q.x;
        p_y[slot][tid] = q.y;
        p_c[slot][tid] = q.c;
        p_wr[slot][tid] = q.wre;
        p_wi[slot][tid] = q.wim;
    };
    auto generate = [&](int64_t c) {
        const int slot = (int)(c & 1);
        if (full)
            fr_generate_pairs<T, SPLIT, true, true>(L[slot], p_x[slot], p_y[slot], p_c[slot], p_wr[slot], p_wi[slot],
                                              tid, i0, j0, a);
        else
            fr_generate_pairs<T, SPLIT, false, true>(L[slot], p_x[slot], p_y[slot], p_c[slot], p_wr[slot], p_wi[slot],
                                               tid, i0, j0, a);
    };
    FrPoint nq{};
    if (tid < K) {
        if (nch > 0) stage_points(0, a.pts[tid]);
        if (nch > 1) stage_points(1, a.pts[K + tid]);
        if (nch > 2) nq = a.pts[2 * K + tid];
    }
    __syncthreads();
    if (nch > 0) generate(0);
    __syncthreads();
    for (int64_t c = 0; c < nch; ++c) {
        if (c + 1 < nch) generate(c + 1);
        fr_contract<T, SPLIT>(L[c & 1], wr, wc, lane, cr, ci);
        // slot c & 1 held chunk c's points, read by generate(c) one step ago
        if (tid < K && c + 2 < nch) {
            stage_points((int)(c & 1), nq);
            if (c + 3 < nch) nq = a.pts[(c + 3) * K + tid];
        }
        // the plane ends with its last chunk: flush, then the next plane's end
        // (a load once per plane, not per chunk: in the vector memory count it
        // would wait for the point prefetch)
        if (c == pend) {
            fr_flush(cr, ci, pl, wr, wc, lane, i0, j0, a);
            ++pl;
            if (pl < a.info[1]) pend = (int64_t)(a.poff[pl + 1] / K) - 1;
        }
        __syncthreads();
    }
}

template <typename T, bool SPLIT>
int launch_fresnel_mfma_planes(oph_ctx* ctx, const PlaneMfmaArgs& m, unsigned tiles) {
    const size_t shm = 2 * sizeof(FrLds<T, SPLIT>);
    static const bool attr = [&] {
        hipFuncSetAttribute(reinterpret_cast<const void*>(&fresnel_mfma_planes<T, SPLIT>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        hipGetLastError();
        return true;
    }();
    (void)attr;
    OPH_LAUNCH(ctx, "fresnel_mfma", (fresnel_mfma_planes<T, SPLIT>), dim3(tiles), dim3(MFMA_THREADS), shm, m);
    return OPH_OK;
}

// Groups the points by plane (device side) and runs the plane-factored
// contraction; *info_out receives the device flag word the general path
// checks (its [0] is nonzero when this form produced the field).
int fresnel_planes(oph_ctx* ctx, int method, int nx, int ny, double pitch, double lambda, const double* dx,
                   const double* dy, const double* dz, const double* da, const double* dp, int64_t n, float2* out,
                   unsigned tiles, const int** info_out) {
    const int64_t cap_chunks = (n + (int64_t)PLANE_MAX * (KC - 1)) / KC + 1;
    char* w0 = static_cast<char*>(oph_scratch(ctx, SLOT_WORK0, 32 * (size_t)n + 256));
    char* tb = static_cast<char*>(oph_scratch(ctx, SLOT_TABLE, 16 * (PLANE_MAX + 1) + 256));
    // plane chirp tables: at most PLANE_D_MAX entries (the form is off beyond)
    const int64_t dlen = (int64_t)nx + ny;
    const int64_t dent = std::min<int64_t>((n / PLANE_MIN_AVG) * dlen, PLANE_D_MAX);
    float2* dtab = static_cast<float2*>(oph_scratch(ctx, SLOT_TABLE2, sizeof(float2) * (size_t)std::max<int64_t>(dent, 1)));
    if (!dtab) return OPH_ENOMEM;
    FrPoint* gp = static_cast<FrPoint*>(oph_scratch(ctx, SLOT_WORK1, sizeof(FrPoint) * cap_chunks * KC));
    if (!w0 || !tb || !gp) return OPH_ENOMEM;
    auto* kin = reinterpret_cast<unsigned long long*>(w0);
    auto* kout = kin + n;
    auto* iin = reinterpret_cast<unsigned*>(kout + n);
    unsigned* iout = iin + n;
    unsigned* flag = iout + n;
    unsigned* pid = flag + n;
    auto* pc = reinterpret_cast<double*>(tb);
    auto* start = reinterpret_cast<unsigned*>(pc + PLANE_MAX);
    unsigned* poff = start + PLANE_MAX + 1;
    int* info = reinterpret_cast<int*>(poff + PLANE_MAX + 1);
    size_t sort_bytes = 0, scan_bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, sort_bytes, kin, kout, iin, iout, (size_t)n, 0, 64, ctx->stream) !=
            hipSuccess ||
        rocprim::inclusive_scan(nullptr, scan_bytes, flag, pid, (size_t)n, rocprim::plus<unsigned>(), ctx->stream) !=
            hipSuccess)
        return oph_fail(ctx, OPH_EDEVICE, "plane grouping: temporary storage query failed");
    size_t tmp_bytes = std::max(sort_bytes, scan_bytes);
    void* tmp = oph_scratch(ctx, SLOT_AUX1, tmp_bytes);
    if (!tmp) return OPH_ENOMEM;
    const unsigned gb = (unsigned)((n + 255) / 256);
    OPH_LAUNCH(ctx, "fresnel_group", fr_keys, dim3(gb), dim3(256), 0, dz, n, kin, iin);
    if (rocprim::radix_sort_pairs(tmp, sort_bytes, kin, kout, iin, iout, (size_t)n, 0, 64, ctx->stream) !=
        hipSuccess)
        return oph_fail(ctx, OPH_EDEVICE, "plane grouping: sort failed");
    OPH_LAUNCH(ctx, "fresnel_group", fr_flags, dim3(gb), dim3(256), 0, kout, n, flag);
    if (rocprim::inclusive_scan(tmp, scan_bytes, flag, pid, (size_t)n, rocprim::plus<unsigned>(), ctx->stream) !=
        hipSuccess)
        return oph_fail(ctx, OPH_EDEVICE, "plane grouping: scan failed");
    OPH_LAUNCH(ctx, "fresnel_group", fr_starts, dim3(gb), dim3(256), 0, kout, flag, pid, n, lambda, start, pc,
               info);
    OPH_LAUNCH(ctx, "fresnel_group", fr_plan, dim3(1), dim3(1024), 0, n, start, poff, info, cap_chunks, dlen);
    const int64_t fill_n = std::max<int64_t>(n, (int64_t)PLANE_MAX * KC);
    OPH_LAUNCH(ctx, "fresnel_group", fr_fill, dim3((unsigned)((fill_n + 255) / 256)), dim3(256), 0, dx, dy, dz, da,
               dp, n, lambda, iout, pid, start, poff, pc, info, gp);
    // rows' tables first, then the columns': np (ny + nx) <= dent entries
    float2* dr = dtab;
    float2* dcl = dtab + (size_t)(dent / dlen) * ny;
    OPH_LAUNCH(ctx, "fresnel_group", fr_chirps, dim3((unsigned)((dent + 255) / 256)), dim3(256), 0, pc, info, nx, ny,
               pitch, dr, dcl);
    PlaneMfmaArgs m{};
    m.pts = gp;
    m.dr = dr;
    m.dc = dcl;
    m.poff = poff;
    m.pc = pc;
    m.info = info;
    m.nx = nx;
    m.ny = ny;
    m.p = pitch;
    m.out = out;
    int rc;
    switch (method) {
        case OPH_FRESNEL_MFMA_BF16: rc = launch_fresnel_mfma_planes<__bf16, false>(ctx, m, tiles); break;
        case OPH_FRESNEL_MFMA_BF16_SPLIT: rc = launch_fresnel_mfma_planes<__bf16, true>(ctx, m, tiles); break;
        case OPH_FRESNEL_MFMA_FP16: rc = launch_fresnel_mfma_planes<_Float16, false>(ctx, m, tiles); break;
        default: rc = launch_fresnel_mfma_planes<_Float16, true>(ctx, m, tiles); break;
    }
    if (rc) return rc;
    *info_out = info;
    return OPH_OK;
}

}  // namespace

extern "C" int oph_pc_fresnel(oph_ctx* ctx, int nx, int ny, double pitch, double lambda, const oph_points* pts,
                              int method, oph_c64* out, int flags) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!pts || !out) return oph_fail(ctx, OPH_EINVAL, "null argument");
    if (nx <= 0 || ny <= 0 || nx > 65536 || ny > 65536) return oph_fail(ctx, OPH_EINVAL, "bad hologram size");
    if (!(pitch > 0) || !(lambda > 0)) return oph_fail(ctx, OPH_EINVAL, "bad pitch or wavelength");
    if (method < OPH_FRESNEL_VALU || method > OPH_FRESNEL_MFMA_FP16_SPLIT)
        return oph_fail(ctx, OPH_EINVAL, "unknown method %d", method);
    if (pts->n < 0 || (pts->n > 0 && (!pts->x || !pts->y || !pts->z)))
        return oph_fail(ctx, OPH_EINVAL, "bad point arrays");
    if ((flags & OPH_ASYNC) && !(flags & OPH_DEVICE_PTR))
        return oph_fail(ctx, OPH_EINVAL, "OPH_ASYNC needs device pointers");
    const int64_t n = pts->n;
    const int64_t npad = ((n + KC - 1) / KC) * KC;
    const size_t pb = (size_t)(n > 0 ? n : 1) * sizeof(double);
    const int sf = n > 0 ? flags : (flags | OPH_DEVICE_PTR);
    if (!(flags & OPH_DEVICE_PTR) && !oph_scratch(ctx, SLOT_PTS, 5 * pb)) return OPH_ENOMEM;
    const double* dx = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->x, pb, sf, 0));
    const double* dy = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->y, pb, sf, pb));
    const double* dz = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->z, pb, sf, 2 * pb));
    const double* da = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->amp, pb, sf, 3 * pb));
    const double* dp = static_cast<const double*>(oph_stage_in(ctx, SLOT_PTS, pts->phase, pb, sf, 4 * pb));
    if (n > 0 && (!dx || !dy || !dz)) return OPH_ENOMEM;
    FrPoint* fp = static_cast<FrPoint*>(oph_scratch(ctx, SLOT_AUX0, sizeof(FrPoint) * (npad > 0 ? npad : 1)));
    if (!fp) return OPH_ENOMEM;
    const size_t npix = (size_t)nx * ny;
    float2* o = reinterpret_cast<float2*>(out);
    if (!(flags & OPH_DEVICE_PTR)) {
        o = static_cast<float2*>(oph_scratch(ctx, SLOT_OUT, npix * sizeof(float2)));
        if (!o) return OPH_ENOMEM;
    }
    const unsigned tiles = (unsigned)(((nx + MC - 1) / MC) * ((ny + MR - 1) / MR));
    // MFMA forms: the plane-factored contraction when the points lie on few
    // planes (decided on the device; OPH_FRESNEL_PLANES=0 disables it), the
    // general one otherwise -- both are launched, the one not chosen returns
    const int* skip = nullptr;
    const char* pe = getenv("OPH_FRESNEL_PLANES");
    if (method != OPH_FRESNEL_VALU && n >= PLANE_MIN_AVG && n < ((int64_t)1 << 31) && !(pe && pe[0] == '0')) {
        rc = fresnel_planes(ctx, method, nx, ny, pitch, lambda, dx, dy, dz, da, dp, n, o, tiles, &skip);
        if (rc) return rc;
    }
    if (npad > 0)
        OPH_LAUNCH(ctx, "fresnel_prep", fresnel_prep, dim3((unsigned)((npad + 255) / 256)), dim3(256), 0, dx, dy,
                   dz, da, dp, n, npad, lambda, fp, skip);
    if (method == OPH_FRESNEL_VALU) {
        const int gx = (nx + 63) / 64, gy = (ny + 63) / 64;
        int64_t chunks = (4LL * ctx->cu_count + (int64_t)gx * gy - 1) / ((int64_t)gx * gy);
        const int64_t maxc = (n + 1023) / 1024;
        if (chunks > maxc) chunks = maxc;
        if (chunks > 16) chunks = 16;
        if (chunks < 1) chunks = 1;
        ValuArgs v{};
        v.pts = fp;
        v.n = n;
        v.chunk = n > 0 ? (n + chunks - 1) / chunks : 1;
        v.nx = nx;
        v.ny = ny;
        v.p = pitch;
        v.partial = static_cast<float2*>(oph_scratch(ctx, SLOT_WORK0, chunks * npix * sizeof(float2)));
        if (!v.partial) return OPH_ENOMEM;
        OPH_LAUNCH(ctx, "fresnel_valu", fresnel_valu, dim3(gx, gy, (unsigned)chunks), dim3(16, 16), 0, v);
        OPH_LAUNCH(ctx, "fresnel_sum_chunks", fresnel_sum_chunks, dim3((unsigned)((npix + 255) / 256)), dim3(256),
                   0, v.partial, (int)chunks, (int64_t)npix, o);
    } else {
        MfmaArgs m{};
        m.pts = fp;
        m.npad = npad;
        m.nx = nx;
        m.ny = ny;
        m.p = pitch;
        m.out = o;
        m.skip = skip;
        switch (method) {
            case OPH_FRESNEL_MFMA_BF16: rc = launch_fresnel_mfma<__bf16, false>(ctx, m, tiles); break;
            case OPH_FRESNEL_MFMA_BF16_SPLIT: rc = launch_fresnel_mfma<__bf16, true>(ctx, m, tiles); break;
            case OPH_FRESNEL_MFMA_FP16: rc = launch_fresnel_mfma<_Float16, false>(ctx, m, tiles); break;
            default: rc = launch_fresnel_mfma<_Float16, true>(ctx, m, tiles); break;
        }
        if (rc) return rc;
    }
    if (!(flags & OPH_DEVICE_PTR))
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(out, o, npix * sizeof(float2), hipMemcpyDeviceToHost, ctx->stream));
    return oph_finish(ctx, flags);
}
