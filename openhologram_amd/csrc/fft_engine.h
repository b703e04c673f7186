// fft_engine.h -- LDS-resident Stockham FFT building blocks for gfx950.
//
// A line of L = 16^a * r complex64 values (r in {1, 2, 4, 8}) sits in LDS in
// natural order.  L/16 threads transform it: a radix-16 stage per factor of
// 16 plus one radix-r stage, each stage reading its R inputs at stride L/R,
// applying twiddles from a table tw[m] = exp(-2 pi i m / L) computed in FP64,
// doing the radix-R DFT in registers and writing back in Stockham (self-
// sorting) order, so the output is again in natural order.  Stages are
// separated by workgroup barriers and work in place (all reads of a stage
// complete before its writes).
//
// SIGN = -1 is the forward transform (exp(-i ...), FFTW_FORWARD / cuFFT
// FORWARD), SIGN = +1 the unnormalised inverse.
#pragma once

#include "oph_internal.h"

namespace fft {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
    return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}
// multiply by SIGN*i:  (x + iy) * (s i) = -s y + i s x
template <int SIGN>
__device__ __forceinline__ float2 mul_si(float2 a) {
    return SIGN < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int SIGN>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
    const float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <int SIGN>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
    const float2 s02 = cadd(a0, a2), d02 = csub(a0, a2);
    const float2 s13 = cadd(a1, a3), d13 = mul_si<SIGN>(csub(a1, a3));
    a0 = cadd(s02, s13);
    a2 = csub(s02, s13);
    a1 = cadd(d02, d13);
    a3 = csub(d02, d13);
}

template <int R, int SIGN>
struct Dft;

template <int SIGN>
struct Dft<1, SIGN> {
    __device__ static void run(float2*) {}
};
template <int SIGN>
struct Dft<2, SIGN> {
    __device__ static void run(float2* v) { dft2<SIGN>(v[0], v[1]); }
};
template <int SIGN>
struct Dft<4, SIGN> {
    __device__ static void run(float2* v) { dft4<SIGN>(v[0], v[1], v[2], v[3]); }
};
template <int SIGN>
struct Dft<8, SIGN> {
    __device__ static void run(float2* v) {
        // radix 2 x 4: v[n], n = 4 n1 + n2?  use decimation: even/odd split
        constexpr float h = 0.70710678118654752f;
        dft4<SIGN>(v[0], v[2], v[4], v[6]);
        dft4<SIGN>(v[1], v[3], v[5], v[7]);
        // twiddle odd part by W8^k, k = 0..3 (W8 = exp(SIGN i 2pi/8))
        v[3] = make_float2(h * (v[3].x - SIGN * v[3].y), h * (v[3].y + SIGN * v[3].x));
        v[5] = mul_si<SIGN>(v[5]);
        v[7] = make_float2(h * (-v[7].x - SIGN * v[7].y), h * (-v[7].y + SIGN * v[7].x));
        float2 o[8];
        o[0] = cadd(v[0], v[1]);
        o[4] = csub(v[0], v[1]);
        o[1] = cadd(v[2], v[3]);
        o[5] = csub(v[2], v[3]);
        o[2] = cadd(v[4], v[5]);
        o[6] = csub(v[4], v[5]);
        o[3] = cadd(v[6], v[7]);
        o[7] = csub(v[6], v[7]);
#pragma unroll
        for (int i = 0; i < 8; ++i) v[i] = o[i];
    }
};
template <int SIGN>
struct Dft<16, SIGN> {
    __device__ static void run(float2* v) {
        // 4 x 4: X[k1 + 4 k2] = sum_{n2} W4^{n2 k2} W16^{n2 k1} sum_{n1} W4^{n1 k1} x[4 n1 + n2]
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) dft4<SIGN>(v[n2], v[n2 + 4], v[n2 + 8], v[n2 + 12]);
        // now v[n2 + 4 k1] holds inner DFT k1 of column n2; twiddle W16^{n2 k1}
        constexpr float c1 = 0.92387953251128676f, s1 = 0.38268343236508977f, h = 0.70710678118654752f;
        const float2 w1 = make_float2(c1, SIGN * s1), w2 = make_float2(h, SIGN * h),
                     w3 = make_float2(s1, SIGN * c1), w6 = make_float2(-h, SIGN * h),
                     w9 = make_float2(-c1, -SIGN * s1);
        v[1 + 4] = cmul(v[1 + 4], w1);
        v[1 + 8] = cmul(v[1 + 8], w2);
        v[1 + 12] = cmul(v[1 + 12], w3);
        v[2 + 4] = cmul(v[2 + 4], w2);
        v[2 + 8] = mul_si<SIGN>(v[2 + 8]);
        v[2 + 12] = cmul(v[2 + 12], w6);
        v[3 + 4] = cmul(v[3 + 4], w3);
        v[3 + 8] = cmul(v[3 + 8], w6);
        v[3 + 12] = cmul(v[3 + 12], w9);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) dft4<SIGN>(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
        // v[4 k1 + k2] holds X[k1 + 4 k2]; transpose into natural order
        float2 o[16];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) o[k1 + 4 * k2] = v[4 * k1 + k2];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = o[i];
    }
};

// Radix of the last (non-16) stage of a length-L transform.
__host__ __device__ constexpr int last_radix(int L) {
    int r = L;
    while (r >= 16 && r % 16 == 0) r /= 16;
    return r;  // 1, 2, 4 or 8 for powers of two
}
__host__ __device__ constexpr int n16(int L) {
    int n = 0;
    while (L >= 16 && L % 16 == 0) { L /= 16; ++n; }
    return n;
}

// LDS index with one float2 of padding per 16 elements (spreads the strided
// Stockham writes over the banks).
__device__ __forceinline__ int pad16(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int padded_len(int L) { return L + L / 16; }

// Twiddles of stage s (Ns = product of the earlier radices, radix R) are
// stored as tws[(r-1) * Ns + k] = exp(-2 pi i k r / (Ns R)), k < Ns,
// r = 1..R-1, so that the lanes of a wave, which hold consecutive k, read
// consecutive entries.  The tables of all stages of one length are
// concatenated; stage 0 (Ns = 1) has none.

// One Stockham stage of radix R over a line in LDS (padded indexing), with the
// stage's twiddles already in registers (w[p][r], r >= 1; unused for Ns = 1).
template <int L, int R, int SIGN, int TPL>
__device__ __forceinline__ void stage(float2* buf, const float2 (&w)[L / R / TPL][R], int Ns, int t) {
    constexpr int NB = L / R;
    constexpr int PER = NB / TPL;  // butterflies per thread (TPL divides NB)
    float2 v[PER][R];
    // padded index of j + r*NB is pad16(j) + r*(NB + NB/16) when 16 | NB:
    // the per-r part folds into the LDS instruction's immediate offset
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int j = t + p * TPL;
        if constexpr (NB % 16 == 0) {
            const float2* src = buf + pad16(j);
#pragma unroll
            for (int r = 0; r < R; ++r) v[p][r] = src[r * (NB + NB / 16)];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) v[p][r] = buf[pad16(j + r * NB)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int j = t + p * TPL;
        const int k = j & (Ns - 1);
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r)
                v[p][r] = SIGN < 0 ? cmul(v[p][r], w[p][r]) : cmulc(v[p][r], w[p][r]);
        }
        Dft<R, SIGN>::run(v[p]);
        const int base = (j - k) * R + k;
        if (Ns == 1) {
            // base = R j: pad16(R j + r) = pad16(R j) + r for r < R <= 16
            float2* dst = buf + pad16(base);
#pragma unroll
            for (int r = 0; r < R; ++r) dst[r] = v[p][r];
        } else if ((Ns & 15) == 0) {
            float2* dst = buf + pad16(base);
#pragma unroll
            for (int r = 0; r < R; ++r) dst[r * (Ns + Ns / 16)] = v[p][r];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) buf[pad16(base + r * Ns)] = v[p][r];
        }
    }
    __syncthreads();
}

// Element i = t + m*TPL of a line, m = 0..15, padded: pad16(t) + m*(TPL + TPL/16)
// when 16 | TPL (used by the kernels' global<->LDS copies).
template <int L>
__device__ __forceinline__ int lane_pad(int t, int m) {
    constexpr int TPL = L / 16;
    if constexpr (TPL % 16 == 0) return pad16(t) + m * (TPL + TPL / 16);
    else return pad16(t + m * TPL);
}

// Twiddles of every stage of one thread's butterflies, held in registers for
// the whole kernel: loaded once, before the data, and shared by a forward and
// an inverse transform of the same length (the inverse conjugates).
// Stage tables are laid out as tws[(r-1) * Ns + k] = exp(-2 pi i k r/(Ns R))
// and concatenated over the stages (oph_twiddles).
template <int L>
struct Twiddles {
    static constexpr int TPL = L / 16;
    static constexpr int A = n16(L);
    static constexpr int RL = last_radix(L);
    static constexpr int PL = RL > 1 ? (L / RL) / TPL : 1;   // butterflies per thread, last stage
    float2 w16[A > 1 ? A - 1 : 1][1][16];                   // radix-16 stages s = 1..A-1
    float2 wl[PL][RL > 1 ? RL : 1];                           // last radix-RL stage
    float2 w0[1][16];                                          // stage 0: unused

    __device__ __forceinline__ void load(const float2* __restrict__ tw, int t) {
        int Ns = 16, off = 0;
#pragma unroll
        for (int s = 1; s < A; ++s) {
            const int k = t & (Ns - 1);
#pragma unroll
            for (int r = 1; r < 16; ++r) w16[s - 1][0][r] = tw[off + (r - 1) * Ns + k];
            off += 15 * Ns;
            Ns *= 16;
        }
        if constexpr (RL > 1) {
            // Ns = 16^A here
#pragma unroll
            for (int p = 0; p < PL; ++p) {
                const int k = (t + p * TPL) & (Ns - 1);
#pragma unroll
                for (int r = 1; r < RL; ++r) wl[p][r] = tw[off + (r - 1) * Ns + k];
            }
        }
    }
};

// Full transform of one LDS line by TPL = L/16 threads (t in [0, TPL)).
// Caller syncs before (line loaded) and the last stage ends with a barrier.
template <int L, int SIGN>
__device__ __forceinline__ void line(float2* buf, const Twiddles<L>& T, int t) {
    constexpr int TPL = L / 16;
    constexpr int A = n16(L);
    constexpr int RL = last_radix(L);
    stage<L, 16, SIGN, TPL>(buf, T.w0, 1, t);
    int Ns = 16;
#pragma unroll
    for (int s = 1; s < A; ++s) {
        stage<L, 16, SIGN, TPL>(buf, T.w16[s - 1], Ns, t);
        Ns *= 16;
    }
    if constexpr (RL > 1) stage<L, RL, SIGN, TPL>(buf, T.wl, Ns, t);
}

// Twiddles of one stage straight from the table (L1/L2-resident), for
// kernels whose registers are needed elsewhere.  Stage tables are
// concatenated from Ns = 16 on: radix-16 stage s (Ns = 16^s) starts at
// sum_{s' < s} 15 * 16^s'; the last radix-RL stage follows them.
template <int L, int R, int TPL>
__device__ __forceinline__ void load_stage_tw(const float2* __restrict__ tw, int off, int Ns, int t,
                                              float2 (&w)[L / R / TPL][R]) {
#pragma unroll
    for (int p = 0; p < L / R / TPL; ++p) {
        const int k = (t + p * TPL) & (Ns - 1);
#pragma unroll
        for (int r = 1; r < R; ++r) w[p][r] = tw[off + (r - 1) * Ns + k];
    }
}

// line() with each stage's twiddles loaded from the table just before it.
template <int L, int SIGN>
__device__ __forceinline__ void line_tw(float2* buf, const float2* __restrict__ tw, int t) {
    constexpr int TPL = L / 16;
    constexpr int A = n16(L);
    constexpr int RL = last_radix(L);
    float2 w0[1][16];
    stage<L, 16, SIGN, TPL>(buf, w0, 1, t);
    int Ns = 16, off = 0;
#pragma unroll
    for (int s = 1; s < A; ++s) {
        float2 w[1][16];
        load_stage_tw<L, 16, TPL>(tw, off, Ns, t, w);
        stage<L, 16, SIGN, TPL>(buf, w, Ns, t);
        off += 15 * Ns;
        Ns *= 16;
    }
    if constexpr (RL > 1) {
        float2 w[L / RL / TPL][RL];
        load_stage_tw<L, RL, TPL>(tw, off, Ns, t, w);
        stage<L, RL, SIGN, TPL>(buf, w, Ns, t);
    }
}

// First and last stages with the thread's elements in registers, for lines
// of L = 16^a (one radix-16 butterfly per thread and stage, TPL = L/16).
// Thread t's stage-0 inputs are elements t + r L/16 (r = 0..15) and its
// last-stage outputs are the same elements, so a kernel can take the line
// from global memory straight into the first butterfly and write the last
// butterfly straight back, without staging the line in LDS.
template <int L>
__host__ __device__ constexpr bool regs_fusable() { return n16(L) >= 2 && last_radix(L) == 1; }

// Stage 0 (Ns = 1) on v (elements t + r L/16), result written to the line.
template <int L, int SIGN>
__device__ __forceinline__ void stage0_from_regs(float2 (&v)[16], float2* buf, int t) {
    static_assert(regs_fusable<L>(), "length must be a power of 16");
    Dft<16, SIGN>::run(v);
    float2* dst = buf + pad16(16 * t);
#pragma unroll
    for (int r = 0; r < 16; ++r) dst[r] = v[r];
}

// Middle stages s = 1 .. a-2 of line_tw (twiddles from the table).
template <int L, int SIGN>
__device__ __forceinline__ void middle_stages_tw(float2* buf, const float2* __restrict__ tw, int t) {
    constexpr int TPL = L / 16;
    constexpr int A = n16(L);
    int Ns = 16, off = 0;
#pragma unroll
    for (int s = 1; s < A - 1; ++s) {
        float2 w[1][16];
        load_stage_tw<L, 16, TPL>(tw, off, Ns, t, w);
        stage<L, 16, SIGN, TPL>(buf, w, Ns, t);
        off += 15 * Ns;
        Ns *= 16;
    }
}

// Last stage with the thread's 15 twiddles (k = t, Ns = L/16) in wl[1..15].
template <int L, int SIGN>
__device__ __forceinline__ void last_stage_to_regs_w(const float2* buf, const float2 (&wl)[16], int t,
                                                     float2 (&v)[16]) {
    static_assert(regs_fusable<L>(), "length must be a power of 16");
    constexpr int TPL = L / 16;
    const float2* src = buf + pad16(t);
#pragma unroll
    for (int r = 0; r < 16; ++r) v[r] = src[r * (TPL + TPL / 16)];
#pragma unroll
    for (int r = 1; r < 16; ++r) v[r] = SIGN < 0 ? cmul(v[r], wl[r]) : cmulc(v[r], wl[r]);
    Dft<16, SIGN>::run(v);
}

// The 15 last-stage twiddles W^(r t) (W = exp(-2 pi i / L), r = 1..15) from
// w1 = W^t and w4 = W^(4t): products at most three deep (relative error
// ~3e-7), for kernels that cannot hold the 15 values in registers.
__device__ __forceinline__ void twiddle_powers(float2 w1, float2 w4, float2 (&wl)[16]) {
    wl[1] = w1;
    wl[2] = cmul(w1, w1);
    wl[3] = cmul(wl[2], w1);
    wl[4] = w4;
    wl[8] = cmul(w4, w4);
    wl[12] = cmul(wl[8], w4);
#pragma unroll
    for (int q = 4; q < 16; q += 4) {
        wl[q + 1] = cmul(wl[q], wl[1]);
        wl[q + 2] = cmul(wl[q], wl[2]);
        wl[q + 3] = cmul(wl[q], wl[3]);
    }
}

// Offset of the last radix-16 stage's table (Ns = L/16) in a line's table.
template <int L>
__host__ __device__ constexpr int last_stage_tw_offset() {
    int off = 0, Ns = 16;
    for (int s = 1; s < n16(L) - 1; ++s) {
        off += 15 * Ns;
        Ns *= 16;
    }
    return off;
}

// Radix-8 variant: L/8 threads per line, each holding 8 elements
// (t + m L/8, m = 0..7), radix-8 stages plus one radix-2/4 stage, twiddles
// of each stage loaded from the radix-8 table just before it.  Half the
// per-thread state of the radix-16 line, for kernels that keep a lot of
// other data in registers.
__host__ __device__ constexpr int n8(int L) {
    int n = 0;
    while (L >= 8 && L % 8 == 0) { L /= 8; ++n; }
    return n;
}
__host__ __device__ constexpr int last_radix8(int L) {
    while (L >= 8 && L % 8 == 0) L /= 8;
    return L;
}

// Length of the radix-8 twiddle table of an L-point line (all stages).
__host__ __device__ constexpr int table_len8(int L) {
    int n = 0, Ns = 1, rem = L;
    while (rem > 1) {
        const int R = rem % 8 == 0 ? 8 : rem;
        if (Ns > 1) n += (R - 1) * Ns;
        Ns *= R;
        rem /= R;
    }
    return n;
}

template <int L, int SIGN>
__device__ __forceinline__ void line_r8(float2* buf, const float2* __restrict__ tw, int t) {
    constexpr int TPL = L / 8;
    constexpr int A = n8(L);
    constexpr int RL = last_radix8(L);
    float2 w0[1][8];
    stage<L, 8, SIGN, TPL>(buf, w0, 1, t);
    int Ns = 8, off = 0;
#pragma unroll
    for (int s = 1; s < A; ++s) {
        float2 w[1][8];
        load_stage_tw<L, 8, TPL>(tw, off, Ns, t, w);
        stage<L, 8, SIGN, TPL>(buf, w, Ns, t);
        off += 7 * Ns;
        Ns *= 8;
    }
    if constexpr (RL > 1) {
        float2 w[L / RL / TPL][RL];
        load_stage_tw<L, RL, TPL>(tw, off, Ns, t, w);
        stage<L, RL, SIGN, TPL>(buf, w, Ns, t);
    }
}

// line_r8 with the thread's elements in registers at both ends: on entry
// v[m] is input element t + m L/8, on return output element t + m L/8.
// Stage 0 (Ns = 1) takes v directly; for L = 8^a the last stage (Ns = L/8,
// butterfly j = t) reads and writes exactly those elements, so its outputs
// stay in v; otherwise the last stage goes through the line and v is read
// back.  The caller syncs before the line is written again (other threads
// may still read the last stage's inputs).
template <int L, int SIGN>
__device__ __forceinline__ void line_r8_regs(float2 (&v)[8], float2* buf, const float2* __restrict__ tw, int t) {
    static_assert(L >= 64, "at least two radix-8 stages");
    constexpr int TPL = L / 8;
    constexpr int A = n8(L);
    constexpr int RL = last_radix8(L);
    Dft<8, SIGN>::run(v);
    {
        float2* dst = buf + pad16(8 * t);
#pragma unroll
        for (int r = 0; r < 8; ++r) dst[r] = v[r];
    }
    __syncthreads();
    int Ns = 8, off = 0;
    constexpr int NMID = RL == 1 ? A - 2 : A - 1;
#pragma unroll
    for (int s = 0; s < NMID; ++s) {
        float2 w[1][8];
        load_stage_tw<L, 8, TPL>(tw, off, Ns, t, w);
        stage<L, 8, SIGN, TPL>(buf, w, Ns, t);
        off += 7 * Ns;
        Ns *= 8;
    }
    if constexpr (RL == 1) {
#pragma unroll
        for (int r = 0; r < 8; ++r) v[r] = buf[pad16(t + r * TPL)];
#pragma unroll
        for (int r = 1; r < 8; ++r) {
            const float2 w = tw[off + (r - 1) * TPL + t];
            v[r] = SIGN < 0 ? cmul(v[r], w) : cmulc(v[r], w);
        }
        Dft<8, SIGN>::run(v);
    } else {
        float2 w[L / RL / TPL][RL];
        load_stage_tw<L, RL, TPL>(tw, off, Ns, t, w);
        stage<L, RL, SIGN, TPL>(buf, w, Ns, t);
#pragma unroll
        for (int m = 0; m < 8; ++m) v[m] = buf[pad16(t + m * TPL)];
    }
}

// Convenience: load the twiddles and transform.
template <int L, int SIGN>
__device__ __forceinline__ void line(float2* buf, const float2* __restrict__ tw, int t) {
    Twiddles<L> T;
    T.load(tw, t);
    line<L, SIGN>(buf, T, t);
}

}  // namespace fft

// Device twiddle tables of all stages of a length-L transform (layout above,
// FP64-computed, rounded to float), cached per length in the context; radix
// 16 (line, line_tw) or 8 (line_r8).  Returns nullptr on failure (error
// recorded).
const float2* oph_twiddles(oph_ctx* ctx, int L, int radix = 16);
