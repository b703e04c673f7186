// fft_mixed.h -- LDS-resident Stockham FFT for lengths N = 2^a 3^b 5^c on
// gfx950 (the SLM grids that are not powers of two, e.g. 11520 x 2160).
//
// A line of N complex64 values sits in LDS in natural order (fft::pad16
// indexing).  NT threads transform it, stage by stage over a compile-time
// radix plan: radix 16 while 16 divides what is left, then one radix 8, 4
// or 2 for the rest of the power of two, then the 5s and the 3s.  A stage of
// radix R after radices of product Ns has N/R butterflies; butterfly j reads
// j + r N/R (r < R), multiplies by exp(SIGN i 2 pi r k/(Ns R)), k = j mod Ns,
// does the R-point DFT in registers and writes (j - k) R + k + r Ns
// (self-sorting, output in natural order).  Each thread keeps all of its
// butterflies' inputs in registers across the stage's barrier, so stages
// work in place.  Twiddles come from a per-length table (oph_twiddles_mixed)
// laid out as the power-of-two engine's: stage s >= 1 holds
// tws[(r-1) Ns + k] = exp(-2 pi i k r/(Ns R)), stages concatenated.
#pragma once

#include "fft_engine.h"

namespace fft {

// 3-point DFT: X1,2 = x0 - (x1 + x2)/2 +- SIGN i (sqrt 3/2)(x1 - x2)
template <int SIGN>
struct Dft<3, SIGN> {
    __device__ static void run(float2* v) {
        constexpr float h = 0.86602540378443865f;
        const float2 t1 = cadd(v[1], v[2]);
        const float2 t2 = make_float2(fmaf(-0.5f, t1.x, v[0].x), fmaf(-0.5f, t1.y, v[0].y));
        const float2 d = csub(v[1], v[2]);
        const float2 t3 = mul_si<SIGN>(make_float2(h * d.x, h * d.y));
        v[0] = cadd(v[0], t1);
        v[1] = cadd(t2, t3);
        v[2] = csub(t2, t3);
    }
};

// 5-point DFT with c_k = cos(2 pi k/5), s_k = sin(2 pi k/5):
//   X0 = x0 + a1 + a2 (a1 = x1 + x4, a2 = x2 + x3, b1 = x1 - x4, b2 = x2 - x3)
//   X1,4 = x0 + c1 a1 + c2 a2 +- SIGN i (s1 b1 + s2 b2)
//   X2,3 = x0 + c2 a1 + c1 a2 +- SIGN i (s2 b1 - s1 b2)
template <int SIGN>
struct Dft<5, SIGN> {
    __device__ static void run(float2* v) {
        constexpr float c1 = 0.30901699437494742f, c2 = -0.80901699437494742f;
        constexpr float s1 = 0.95105651629515357f, s2 = 0.58778525229247313f;
        const float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        const float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        const float2 m1 = make_float2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
        const float2 m2 = make_float2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
        const float2 n1 = mul_si<SIGN>(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
        const float2 n2 = mul_si<SIGN>(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
        v[0] = make_float2(v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y);
        v[1] = cadd(m1, n1);
        v[4] = csub(m1, n1);
        v[2] = cadd(m2, n2);
        v[3] = csub(m2, n2);
    }
};

}  // namespace fft

namespace fftm {

struct Plan {
    int n;       // number of stages (-1: N has a prime factor other than 2, 3, 5)
    int r[24];   // radices in order
};

__host__ __device__ constexpr Plan plan(int N) {
    Plan p{0, {}};
    int rest = N;
    while (rest % 16 == 0) { p.r[p.n++] = 16; rest /= 16; }
    if (rest % 8 == 0) { p.r[p.n++] = 8; rest /= 8; }
    else if (rest % 4 == 0) { p.r[p.n++] = 4; rest /= 4; }
    else if (rest % 2 == 0) { p.r[p.n++] = 2; rest /= 2; }
    while (rest % 5 == 0) { p.r[p.n++] = 5; rest /= 5; }
    while (rest % 3 == 0) { p.r[p.n++] = 3; rest /= 3; }
    if (rest != 1) p.n = -1;
    return p;
}

// Threads per line: the radix-16 stage's butterfly count rounded up to whole
// waves (every stage then has at most ~6 butterflies per thread), at most
// 512 (a longer line keeps more elements per thread, but 8 waves of up to
// 256 registers each, where 12 would be held to 168).
__host__ __device__ constexpr int threads(int N) {
    const int t = ((N + 15) / 16 + 63) / 64 * 64;
    return t < 512 ? t : 512;
}
// Elements per thread when a thread owns t + m NT, m < ept(N).
__host__ __device__ constexpr int ept(int N) { return (N + threads(N) - 1) / threads(N); }

// Marks every element of w (but w[0]) as needed at this point, after all
// of them were loaded: the loads are issued as one batch and waited for
// once (left to itself the compiler orders load, wait, use per element).
template <int R>
__device__ __forceinline__ void batch_ready(float2 (&w)[R]) {
#pragma unroll
    for (int r = 1; r < R; ++r) asm volatile("" : "+v"(w[r].x), "+v"(w[r].y));
}

// One stage of radix R (Ns = product of the earlier radices) by NT threads;
// tws is the stage's table (unused for Ns = 1).  Starts and ends with the
// line complete in LDS (a barrier separates its reads from its writes and
// ends it).
template <int N, int R, int NS, int NT, int SIGN>
__device__ __forceinline__ void stage(float2* buf, const float2* __restrict__ tws, int t) {
    constexpr int NB = N / R;
    constexpr int PER = (NB + NT - 1) / NT;
    float2 v[PER][R];
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int j = t + p * NT;
        if (PER * NT == NB || j < NB) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[p][r] = buf[fft::pad16(j + r * NB)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int j = t + p * NT;
        if (PER * NT == NB || j < NB) {
            const int k = j % NS;
            if constexpr (NS > 1) {
                // the butterfly's R - 1 twiddles are all requested before the
                // first is used: one wait for the batch, not one per load
                float2 w[R];
#pragma unroll
                for (int r = 1; r < R; ++r) w[r] = tws[(r - 1) * NS + k];
                batch_ready(w);
#pragma unroll
                for (int r = 1; r < R; ++r) v[p][r] = SIGN < 0 ? fft::cmul(v[p][r], w[r]) : fft::cmulc(v[p][r], w[r]);
            }
            fft::Dft<R, SIGN>::run(v[p]);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[fft::pad16(base + r * NS)] = v[p][r];
        }
    }
    __syncthreads();
}

template <int N, int NT, int SIGN, int S, int NS, int OFF>
__device__ __forceinline__ void line_from(float2* buf, const float2* __restrict__ tw, int t) {
    constexpr Plan P = plan(N);
    if constexpr (S < P.n) {
        constexpr int R = P.r[S];
        stage<N, R, NS, NT, SIGN>(buf, tw + OFF, t);
        line_from<N, NT, SIGN, S + 1, NS * R, OFF + (NS > 1 ? (R - 1) * NS : 0)>(buf, tw, t);
    }
}

// Full transform of one LDS line (natural order in and out) by threads
// t < NT of the workgroup; every thread of the workgroup must call it (the
// stages hold barriers).  Caller syncs before (line loaded).
template <int N, int NT, int SIGN>
__device__ __forceinline__ void line(float2* buf, const float2* __restrict__ tw, int t) {
    static_assert(plan(N).n > 0, "length must be 2^a 3^b 5^c");
    line_from<N, NT, SIGN, 0, 1, 0>(buf, tw, t);
}

// LDS elements of a padded line
__host__ __device__ constexpr int padded_len(int N) { return N + N / 16 + 1; }

}  // namespace fftm

// Device twiddle table of the mixed-radix plan of length N (layout above,
// FP64-computed, rounded to float), cached per length in the context.
// Returns nullptr on failure (error recorded).
const float2* oph_twiddles_mixed(oph_ctx* ctx, int N);
