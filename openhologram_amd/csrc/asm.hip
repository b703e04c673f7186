// asm.hip -- angular-spectrum propagation with zero padding, fused into three
// FFT passes on gfx950.
//
// Follows the plumbing of CWO::Diffract (CWO/cwo.cpp:2198-2278: pad N -> pad*N
// centred at (Np/2 - N/2), transform, crop the same window) with the
// CWO_ANGULAR transfer (CWO/cwo_lib.h:66-101) and the open statement of the
// band-limited form (OpenHolo_LFCGH/fn_FresnelPropagation_as.m:17-44):
//   H(fx, fy) = exp(i 2 pi z sqrt(1/lambda^2 - fx^2 - fy^2)), 0 if evanescent
// on FFT-order grids fx = kx/(Npx px), kx in [-Npx/2, Npx/2).
//
// Passes (work buffer Npy x Npx complex64 in HBM):
//   1. rows:    each input row is embedded at column offset ox in an LDS line,
//               transformed, and written to work row oy + r (the other rows
//               of the work buffer are known to be zero and never touched);
//   2. columns: B columns of work rows [oy, oy+ny) are staged in LDS, the
//               column FFT, the transfer multiply and the inverse column FFT
//               run back to back in LDS, and only rows [oy, oy+ny) -- the
//               crop window -- are written back;
//   3. rows:    inverse row transform of work rows [oy, oy+ny), scaled by
//               1/(Npx Npy), columns [ox, ox+nx) written to the output.
// The transfer phase in turns is split as
//   z/lambda * sqrt(1 - u) = z/lambda - z lambda (fx^2+fy^2)/2
//                           - z/lambda (u^2/8 + u^3/16 + 5u^4/128 + 7u^5/256 + ...)
// with u = lambda^2 (fx^2 + fy^2): the constant and the per-axis Fresnel terms
// frac(-z lambda fx^2/2) are formed in FP64 (tables per axis, ~1e5 turns of
// range), the small remainder in FP32.  When the truncated series would be
// off by more than 1e-7 turns at the grid corner the kernel evaluates the
// square root per element in FP64 instead.
#include <algorithm>
#include <cmath>

#include "fft_engine.h"
#include "fft_kernels.h"
#include "pipelines.h"

namespace {

// per-axis table entry: x = frac(-z lambda f^2 / 2) (turns), y = lambda^2 f^2,
// z = (|f| z lambda / (Np p / 2))^2 for the band limit, w = f (cycles/m);
// entries [0, nx) for the x axis, [nx, nx + ny) for the y axis.  The forward
// row pass fills them (the column pass is the first reader), spread over its
// threads, so they cost no launch of their own.
struct TabJob {
    float4* t;
    float2* t2;  // y axis again as (x, y) only: what the column pass reads without a band limit
    int nx, ny;
    double px, py, lambda, z;
};

__device__ void fill_axis_tables(const TabJob& j, int first, int step) {
    for (int k = first; k < j.nx + j.ny; k += step) {
        int kk = k, n = j.nx;
        double pitch = j.px;
        if (k >= j.nx) {
            kk = k - j.nx;
            n = j.ny;
            pitch = j.py;
        }
        const int ks = kk < n / 2 ? kk : kk - n;
        const double f = ks / (n * pitch);
        double tt = -j.z * j.lambda * f * f * 0.5;
        tt -= rint(tt);
        const double bl = fabs(f) * j.z * j.lambda / (n * pitch * 0.5);
        j.t[k] = make_float4((float)tt, (float)(j.lambda * j.lambda * f * f), (float)(bl * bl), (float)f);
        if (k >= j.nx) j.t2[kk] = make_float2((float)tt, (float)(j.lambda * j.lambda * f * f));
    }
}

struct AsmConst {
    float c0, K;          // frac(z/lambda), z/lambda
    double zl;            // z/lambda
    double dfx, dfy;      // frequency steps 1/(Np p)
    double lam2;          // lambda^2
    int npx, npy;
    int band;
};

// exp(i 2 pi (c0 + tx + ty - K poly(u))) or 0; EXACT: z/lambda sqrt(1-u) in FP64
// (ay: the y entry's phase and lambda^2 f^2; BAND: the band-limit term of the
// y entry is read from the full table).  Branch-free: the phase is evaluated
// everywhere and the zero selected afterwards, so a caller's unrolled loop
// is one basic block whose table loads can all be in flight together.
// bz: the row's band-limit term (taby[i].z), loaded by the caller with the
// other row-table entries
template <bool EXACT, bool BAND>
__device__ __forceinline__ float2 transfer(const float4 ax, const float2 ay, float bz, int c, int i,
                                           const AsmConst& k) {
    const float u = ax.y + ay.y;
    bool zero = u >= 1.f;
    if constexpr (BAND) {
        const float s = 1.f - u;
        zero = zero | (ax.z > s) | (bz > s);
    }
    float ph;
    if constexpr (!EXACT) {
        const float p = u * u * fmaf(u, fmaf(u, fmaf(u, 7.f / 256.f, 5.f / 128.f), 1.f / 16.f), 0.125f);
        ph = k.c0 + ax.x + ay.x - k.K * p;
    } else {
        const int kx = c < k.npx / 2 ? c : c - k.npx;
        const int ky = i < k.npy / 2 ? i : i - k.npy;
        const double fx = kx * k.dfx, fy = ky * k.dfy;
        const double uu = k.lam2 * (fx * fx + fy * fy);
        zero = zero | (uu >= 1.0);
        const double t = k.zl * sqrt(fmax(1.0 - uu, 0.0));
        ph = (float)(t - rint(t));
    }
    const float cs = cos_turns(ph), sn = sin_turns(ph);
    return zero ? make_float2(0.f, 0.f) : make_float2(cs, sn);
}

// Multiplies N of the thread's column elements, v[j] = row t + (m0 + j) L/16,
// by the transfer function; the N table entries are loaded before any is used.
template <int L, bool EXACT, bool BAND, int N>
__device__ __forceinline__ void apply_transfer(float2 (&v)[N], const float4 ax, const float2* __restrict__ taby2,
                                               const float4* __restrict__ taby, int c, int t, int m0,
                                               const AsmConst& k) {
    float2 ay[N];
    float bz[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        ay[j] = taby2[t + (m0 + j) * (L / 16)];
        bz[j] = BAND ? taby[t + (m0 + j) * (L / 16)].z : 0.f;
    }
#pragma unroll
    for (int j = 0; j < N; ++j)
        v[j] = fft::cmul(v[j], transfer<EXACT, BAND>(ax, ay[j], bz[j], c, t + (m0 + j) * (L / 16), k));
}

template <int L>
__global__ __launch_bounds__(512) void asm_rows_fwd(const float2* __restrict__ in, int nx, int ny, int ox,
                                                    int oy, float2* __restrict__ work,
                                                    const float2* __restrict__ tw, TabJob tj) {
    fill_axis_tables(tj, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
    constexpr int TPL = L / 16;
    constexpr int LPW = fftk::row_lines(L);
    constexpr int PL = fft::padded_len(L);
    __shared__ float2 lds[LPW * PL];
    const int sub = threadIdx.x / TPL, t = threadIdx.x % TPL;
    const int r = blockIdx.x * LPW + sub;
    float2* buf = lds + sub * PL;
    const bool ok = r < ny;
    const float2* src = in + (size_t)(ok ? r : 0) * nx;
    fft::Twiddles<L> T;
    T.load(tw, t);
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const int c = t + m * TPL - ox;
        buf[fft::lane_pad<L>(t, m)] = (ok && c >= 0 && c < nx) ? src[c] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft::line<L, -1>(buf, T, t);
    if (ok) {
        float2* dst = work + (size_t)(oy + r) * L;
#pragma unroll
        for (int m = 0; m < 16; ++m) dst[t + m * TPL] = buf[fft::lane_pad<L>(t, m)];
    }
}

// Column lengths whose transform is register-fused around the transfer.
template <int L>
__host__ __device__ constexpr bool asm_cols_fused() { return fft::regs_fusable<L>() && fft::n16(L) == 3; }

template <int L, int B, bool EXACT>
__global__ __launch_bounds__(1024) void asm_cols(float2* work, int npx, int ny, int oy,
                                                 const float4* __restrict__ tabx,
                                                 const float4* __restrict__ taby,
                                                 const float2* __restrict__ taby2, AsmConst k,
                                                 const float2* __restrict__ tw) {
    constexpr int TPL = L / 16;
    constexpr int PL = fft::padded_len(L);
    extern __shared__ float2 lds[];
    const int nthr = TPL * B;
    const int g = xcd_remap(blockIdx.x, gridDim.x);
    const int cbase = g * B;
    // 4096-point columns (three radix-16 stages): the middle stage's table
    // (240 entries, k = t mod 16) in LDS before the lines, the last stage's 15
    // twiddles (k = t) read from the table just before each use; other
    // lengths hold every stage's twiddles in registers
    constexpr bool FUSED = asm_cols_fused<L>();
    fft::Twiddles<FUSED ? 16 : L> T;  // (a placeholder when FUSED: not loaded or used)
    float2* tw1 = lds;  // first, so every read of it is one address plus an immediate
    float2* lines = lds + (FUSED ? 256 : 0);
    if constexpr (!FUSED) T.load(tw, threadIdx.x % TPL);
    if constexpr (B % 2 == 0 && (L <= 1024 || FUSED)) {
        // two neighbouring columns per lane (16-byte accesses); npx is a
        // power of two >= B, so every column of the group exists.  All 8
        // loads of a lane are issued before any is used: range-checked
        // buffer loads, rows outside [oy, oy + ny) at an out-of-range offset
        // read zeros (the grid is at most 8192^2 x 8 bytes, offsets fit).
        // (2048 and 8192 keep the loop below: their twiddles leave no room
        // for 8 loads in flight)
        constexpr int PER = (L * B / 2) / (TPL * B);
        const auto rs = oph_buf_rsrc(work, (unsigned)((size_t)(oy + ny) * npx * sizeof(float2)));
        float4 ld[PER];
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int idx = threadIdx.x + j * nthr;
            const int pr = idx % (B / 2), i = idx / (B / 2);
            const bool in = i >= oy && i < oy + ny;
            ld[j] = oph_buf_load_f4(rs, in ? (unsigned)(((size_t)i * npx + cbase + 2 * pr) * sizeof(float2)) : OPH_BUF_OOB);
        }
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int idx = threadIdx.x + j * nthr;
            const int pr = idx % (B / 2), i = idx / (B / 2);
            lines[(2 * pr) * PL + fft::pad16(i)] = make_float2(ld[j].x, ld[j].y);
            lines[(2 * pr + 1) * PL + fft::pad16(i)] = make_float2(ld[j].z, ld[j].w);
        }
    } else if constexpr (B % 2 == 0) {
        for (int idx = threadIdx.x; idx < L * B / 2; idx += nthr) {
            const int pr = idx % (B / 2), i = idx / (B / 2);
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (i >= oy && i < oy + ny)
                v = *reinterpret_cast<const float4*>(work + (size_t)i * npx + cbase + 2 * pr);
            lines[(2 * pr) * PL + fft::pad16(i)] = make_float2(v.x, v.y);
            lines[(2 * pr + 1) * PL + fft::pad16(i)] = make_float2(v.z, v.w);
        }
    } else {
        for (int idx = threadIdx.x; idx < L * B; idx += nthr) {
            const int b = idx % B, i = idx / B;
            const int c = cbase + b;
            const bool in_rows = i >= oy && i < oy + ny;
            lines[b * PL + fft::pad16(i)] =
                (in_rows && c < npx) ? work[(size_t)i * npx + c] : make_float2(0.f, 0.f);
        }
    }
    if constexpr (FUSED) {
        // (after the column loads, so that its loads do not hold them up)
        for (int e = threadIdx.x; e < 15 * 16; e += nthr) tw1[e] = tw[e];
    }
    __syncthreads();
    const int b = threadIdx.x / TPL, t = threadIdx.x % TPL;
    float2* line = lines + b * PL;
    const int c = cbase + b;
    const float4 ax = tabx[c < npx ? c : 0];
    if constexpr (FUSED) {
        // the forward transform's last butterfly leaves elements t + m L/16
        // in registers, which are the inverse's first butterfly's inputs:
        // the transfer is applied in between, without a pass through LDS
        float2 w[1][16];
        fft::stage<L, 16, -1, TPL>(line, w, 1, t);
#pragma unroll
        for (int r = 1; r < 16; ++r) w[0][r] = tw1[(r - 1) * 16 + (t & 15)];
        fft::stage<L, 16, -1, TPL>(line, w, 16, t);
        float2 v[16];
        // (buffer loads: one lane offset, the per-r part in the instruction)
        const auto twl = oph_buf_rsrc(tw + fft::last_stage_tw_offset<L>(), 15 * TPL * sizeof(float2));
#pragma unroll
        for (int r = 1; r < 16; ++r) w[0][r] = oph_buf_load_f2(twl, t * sizeof(float2), (r - 1) * TPL * sizeof(float2));
        fft::last_stage_to_regs_w<L, -1>(line, w[0], t, v);
        if (k.band) apply_transfer<L, EXACT, true>(v, ax, taby2, taby, c, t, 0, k);
        else apply_transfer<L, EXACT, false>(v, ax, taby2, taby, c, t, 0, k);
        __syncthreads();  // every lane has read the line
        fft::stage0_from_regs<L, +1>(v, line, t);
        __syncthreads();
#pragma unroll
        for (int r = 1; r < 16; ++r) w[0][r] = tw1[(r - 1) * 16 + (t & 15)];
        fft::stage<L, 16, +1, TPL>(line, w, 16, t);
#pragma unroll
        for (int r = 1; r < 16; ++r) w[0][r] = oph_buf_load_f2(twl, t * sizeof(float2), (r - 1) * TPL * sizeof(float2));
        fft::stage<L, 16, +1, TPL>(line, w, 16 * 16, t);
    } else {
        fft::line<L, -1>(line, T, t);
        // (a few rows at a time: every stage's twiddles stay in registers)
        constexpr int CH = L > 4096 ? 2 : 4;
#pragma unroll
        for (int m0 = 0; m0 < 16; m0 += CH) {
            float2 v[CH];
#pragma unroll
            for (int j = 0; j < CH; ++j) v[j] = line[fft::lane_pad<L>(t, m0 + j)];
            if (k.band) apply_transfer<L, EXACT, true>(v, ax, taby2, taby, c, t, m0, k);
            else apply_transfer<L, EXACT, false>(v, ax, taby2, taby, c, t, m0, k);
#pragma unroll
            for (int j = 0; j < CH; ++j) line[fft::lane_pad<L>(t, m0 + j)] = v[j];
        }
        __syncthreads();
        fft::line<L, +1>(line, T, t);
    }
    if constexpr (B % 2 == 0) {
        for (int idx = threadIdx.x; idx < ny * B / 2; idx += nthr) {
            const int pr = idx % (B / 2), i = oy + idx / (B / 2);
            const float2 u = lines[(2 * pr) * PL + fft::pad16(i)], v = lines[(2 * pr + 1) * PL + fft::pad16(i)];
            *reinterpret_cast<float4*>(work + (size_t)i * npx + cbase + 2 * pr) = make_float4(u.x, u.y, v.x, v.y);
        }
    } else {
        for (int idx = threadIdx.x; idx < ny * B; idx += nthr) {
            const int bb = idx % B, i = oy + idx / B;
            const int cc = cbase + bb;
            if (cc < npx) work[(size_t)i * npx + cc] = lines[bb * PL + fft::pad16(i)];
        }
    }
}

// 4096-point columns whose data rows are exactly the middle half,
// [1024, 3072) (twofold padding), B columns per step, persistent.  The first
// forward butterfly's inputs t + m 256 with m outside [4, 12) are known zeros
// and only the last inverse butterfly's outputs with m in [4, 12) are kept,
// so neither is read or written in LDS.  Each workgroup walks groups of B
// columns: the next group's rows are fetched into registers while the
// current group is transformed, so the loads' HBM latency overlaps the LDS
// work.  Groups are split by XCD (the hardware places workgroup b on XCD
// b mod 8): XCD x owns groups [x G/8, (x + 1) G/8) and its workgroups take
// consecutive groups at each step, so the row pieces of one 128-byte line
// are read and written through one L2 at about the same time.  npx is a
// multiple of 8 B and the grid a multiple of 8.
constexpr int MID_L = 4096, MID_B = 4;
constexpr int MID_TW = 768;   // LDS entries ahead of the lines: middle-stage table, W^t, W^(4t)

template <int B, bool EXACT, bool BAND>
__global__ __launch_bounds__(B * (MID_L / 16), 4) void asm_cols_mid(
    float2* work, int npx, const float4* __restrict__ tabx, const float4* __restrict__ taby,
    const float2* __restrict__ taby2, AsmConst k, const float2* __restrict__ tw) {
    constexpr int L = MID_L, TPL = L / 16, PL = fft::padded_len(L), NTHR = B * TPL, H = B / 2;
    // float4 row pieces (two columns each) per lane: H per row, rows [L/4, 3L/4)
    constexpr int PER = (L / 2) * H / NTHR;
    extern __shared__ float2 lds[];
    float2* tw1 = lds;        // middle stage's 240 twiddles (k = t mod 16)
    float2* wp = lds + 256;   // [t]: W^t, [TPL + t]: W^(4t), W = exp(-2 pi i/L)
    float2* lines = lds + MID_TW;
    const int per = (npx / B) / 8, steps = gridDim.x / 8;
    const int xcd = blockIdx.x % 8;
    const int gend = (xcd + 1) * per;
    int g = xcd * per + blockIdx.x / 8;
    const auto rs = oph_buf_rsrc(work, (unsigned)((size_t)(3 * L / 4) * npx * sizeof(float2)));
    float4 ld[PER];
    auto fetch = [&](int grp) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int idx = threadIdx.x + j * NTHR;
            const int pr = idx % H, i = L / 4 + idx / H;
            ld[j] = oph_buf_load_f4(rs, (unsigned)(((size_t)i * npx + B * grp + 2 * pr) * sizeof(float2)));
        }
    };
    // the FP64 transfer (EXACT) needs the registers the prefetch would hold:
    // it fetches each group at the top of its step instead
    constexpr bool PREFETCH = !EXACT;
    if (PREFETCH && g < gend) fetch(g);
    constexpr int TWO = fft::last_stage_tw_offset<L>();
    for (int e = threadIdx.x; e < TWO; e += NTHR) tw1[e] = tw[e];
    if constexpr (PREFETCH)
        for (int e = threadIdx.x; e < 2 * TPL; e += NTHR) wp[e] = tw[TWO + (e < TPL ? e : 2 * TPL + e)];
    const int b = threadIdx.x / TPL, t = threadIdx.x % TPL;
    float2* line = lines + b * PL;
    const auto twl = oph_buf_rsrc(tw + TWO, 15 * TPL * sizeof(float2));
    // the last stages' 15 twiddles: with the prefetch in flight they are
    // rebuilt from W^t and W^(4t) in LDS (vector memory counts complete in
    // issue order, so a table load issued behind the prefetch would wait for
    // the next group's rows); without it they are read from the table
    auto last_tw = [&](float2 (&wl)[16], int tz) {
        if constexpr (PREFETCH) {
            fft::twiddle_powers(wp[t], wp[TPL + t], wl);
        } else {
#pragma unroll
            for (int r = 1; r < 16; ++r) wl[r] = oph_buf_load_f2(twl, tz * sizeof(float2), (r - 1) * TPL * sizeof(float2));
        }
    };
    for (; g < gend; g += steps) {
        if constexpr (!PREFETCH) fetch(g);
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int idx = threadIdx.x + j * NTHR;
            const int pr = idx % H, i = L / 4 + idx / H;
            lines[(2 * pr) * PL + fft::pad16(i)] = make_float2(ld[j].x, ld[j].y);
            lines[(2 * pr + 1) * PL + fft::pad16(i)] = make_float2(ld[j].z, ld[j].w);
        }
        __syncthreads();
        // a zero the compiler cannot see through, added to the lane index of
        // the loop-invariant table loads: hoisted out of the loop they would
        // hold ~60 registers across it and spill
        int zero;
        asm volatile("v_mov_b32 %0, 0" : "=v"(zero));
        const int tz = t + zero;
        const int c = B * g + b;
        const float4 ax = tabx[c];
        // forward: stage 0 on the middle rows only (the rest are zeros)
        float2 u[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) u[m] = (m >= 4 && m < 12) ? line[fft::lane_pad<L>(t, m)] : make_float2(0.f, 0.f);
        __syncthreads();  // every lane has read its inputs
        fft::stage0_from_regs<L, -1>(u, line, t);
        __syncthreads();
        float2 w[1][16];
#pragma unroll
        for (int r = 1; r < 16; ++r) w[0][r] = tw1[(r - 1) * 16 + (t & 15)];
        fft::stage<L, 16, -1, TPL>(line, w, 16, t);
        float2 v[16];
        last_tw(w[0], tz);
        // the forward transform's last butterfly leaves elements t + m 256 in
        // registers, the inverse's first butterfly's inputs: the transfer is
        // applied in between, without a pass through LDS
        fft::last_stage_to_regs_w<L, -1>(line, w[0], t, v);
        apply_transfer<L, EXACT, BAND>(v, ax, taby2, taby, c, tz, 0, k);
        __syncthreads();  // every lane has read the line
        fft::stage0_from_regs<L, +1>(v, line, t);
        // the next group's rows: in flight through the inverse transform
        if (PREFETCH && g + steps < gend) fetch(g + steps);
        __syncthreads();
#pragma unroll
        for (int r = 1; r < 16; ++r) w[0][r] = tw1[(r - 1) * 16 + (t & 15)];
        fft::stage<L, 16, +1, TPL>(line, w, 16, t);
        last_tw(w[0], tz);
        // the last butterfly reads and writes only the lane's own elements
        // t + m 256, so no barrier separates the two
        fft::last_stage_to_regs_w<L, +1>(line, w[0], t, v);
#pragma unroll
        for (int m = 4; m < 12; ++m) line[fft::lane_pad<L>(t, m)] = v[m];
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int idx = threadIdx.x + j * NTHR;
            const int pr = idx % H, i = L / 4 + idx / H;
            const float2 a0 = lines[(2 * pr) * PL + fft::pad16(i)], a1 = lines[(2 * pr + 1) * PL + fft::pad16(i)];
            *reinterpret_cast<float4*>(work + (size_t)i * npx + B * g + 2 * pr) = make_float4(a0.x, a0.y, a1.x, a1.y);
        }
        __syncthreads();  // the lines are read out before the next group is stored
    }
}

template <int L>
__global__ __launch_bounds__(512) void asm_rows_inv(const float2* __restrict__ work, int nx, int ny, int ox,
                                                    int oy, float scale, float2* __restrict__ out,
                                                    const float2* __restrict__ tw) {
    constexpr int TPL = L / 16;
    constexpr int LPW = fftk::row_lines(L);
    constexpr int PL = fft::padded_len(L);
    __shared__ float2 lds[LPW * PL];
    const int sub = threadIdx.x / TPL, t = threadIdx.x % TPL;
    const int r = blockIdx.x * LPW + sub;
    float2* buf = lds + sub * PL;
    const bool ok = r < ny;
    const float2* src = work + (size_t)(oy + (ok ? r : 0)) * L;
    fft::Twiddles<L> T;
    T.load(tw, t);
#pragma unroll
    for (int m = 0; m < 16; ++m) buf[fft::lane_pad<L>(t, m)] = ok ? src[t + m * TPL] : make_float2(0.f, 0.f);
    __syncthreads();
    fft::line<L, +1>(buf, T, t);
    if (ok) {
        float2* dst = out + (size_t)r * nx;
        for (int c = t; c < nx; c += TPL) {
            const float2 v = buf[fft::pad16(c + ox)];
            dst[c] = make_float2(v.x * scale, v.y * scale);
        }
    }
}

// Persistent row passes for long rows (one line per workgroup): each
// workgroup walks rows r, r + G, ...; the next row is fetched into registers
// while the current one is transformed in LDS, so the global loads overlap
// the transform (twiddles per stage from the table, to leave registers for
// the prefetch).
template <int L>
__global__ __launch_bounds__(L / 16) void asm_rows_fwd_p(const float2* __restrict__ in, int nx, int ny, int ox,
                                                         int oy, float2* __restrict__ work,
                                                         const float2* __restrict__ tw, TabJob tj) {
    fill_axis_tables(tj, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
    constexpr int TPL = L / 16;
    constexpr int PL = fft::padded_len(L);
    __shared__ float2 buf[PL];
    const int t = threadIdx.x;
    float2 nxt[16];
    int r = blockIdx.x;
#pragma unroll
    for (int m = 0; m < 16; ++m) {
        const int c = t + m * TPL - ox;
        nxt[m] = (r < ny && c >= 0 && c < nx) ? in[(size_t)r * nx + c] : make_float2(0.f, 0.f);
    }
    for (; r < ny; r += gridDim.x) {
#pragma unroll
        for (int m = 0; m < 16; ++m) buf[fft::lane_pad<L>(t, m)] = nxt[m];
        __syncthreads();
        const int rn = r + gridDim.x;
#pragma unroll
        for (int m = 0; m < 16; ++m) {
            const int c = t + m * TPL - ox;
            nxt[m] = (rn < ny && c >= 0 && c < nx) ? in[(size_t)rn * nx + c] : make_float2(0.f, 0.f);
        }
        fft::line_tw<L, -1>(buf, tw, t);
        float2* dst = work + (size_t)(oy + r) * L;
#pragma unroll
        for (int m = 0; m < 16; ++m) dst[t + m * TPL] = buf[fft::lane_pad<L>(t, m)];
        __syncthreads();
    }
}

template <int L>
__global__ __launch_bounds__(L / 16) void asm_rows_inv_p(const float2* __restrict__ work, int nx, int ny, int ox,
                                                         int oy, float scale, float2* __restrict__ out,
                                                         const float2* __restrict__ tw) {
    constexpr int TPL = L / 16;
    constexpr int PL = fft::padded_len(L);
    __shared__ float2 buf[PL];
    const int t = threadIdx.x;
    float2 nxt[16];
    int r = blockIdx.x;
#pragma unroll
    for (int m = 0; m < 16; ++m)
        nxt[m] = r < ny ? work[(size_t)(oy + r) * L + t + m * TPL] : make_float2(0.f, 0.f);
    for (; r < ny; r += gridDim.x) {
#pragma unroll
        for (int m = 0; m < 16; ++m) buf[fft::lane_pad<L>(t, m)] = nxt[m];
        __syncthreads();
        const int rn = r + gridDim.x;
#pragma unroll
        for (int m = 0; m < 16; ++m)
            nxt[m] = rn < ny ? work[(size_t)(oy + rn) * L + t + m * TPL] : make_float2(0.f, 0.f);
        fft::line_tw<L, +1>(buf, tw, t);
        float2* dst = out + (size_t)r * nx;
        for (int c = t; c < nx; c += TPL) {
            const float2 v = buf[fft::pad16(c + ox)];
            dst[c] = make_float2(v.x * scale, v.y * scale);
        }
        __syncthreads();
    }
}

// Persistent row passes for L = 16^a with the first and last transform
// stages in registers: the prefetched row goes straight into the first
// butterfly and the last butterfly straight to global memory (thread t holds
// elements t + m L/16 in both), so the line passes through LDS only between
// stages.  HALF: the data columns [ox, ox + nx) lie inside [L/4, 3L/4)
// (twofold padding), so only m in [4, 12) can be nonzero on input (forward)
// or is needed on output (inverse).  The next row is loaded while the
// current one is transformed: the stage twiddles of a thread do not depend
// on the row and are loaded once (the middle stages' tables into LDS, the
// last stage's 15 per thread into registers), and the row loads and
// stores are range-checked buffer accesses without branches, so the only
// memory wait in the loop is the one for the prefetched row (the wait
// counter is in order; a global twiddle load or a branch around a load
// would make that wait early).  The first row is peeled off the loop, so the
// loop is entered only with a row's loads followed by a row's stores
// outstanding: the wait before each row's first butterfly then covers the
// loads and not the previous row's stores.
template <int L, bool HALF>
__global__ __launch_bounds__(L / 16) void asm_rows_fwd_f(const float2* __restrict__ in, int nx, int ny, int ox,
                                                         int oy, float2* __restrict__ work,
                                                         const float2* __restrict__ tw, TabJob tj) {
    fill_axis_tables(tj, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
    constexpr int TPL = L / 16;
    constexpr int PL = fft::padded_len(L);
    constexpr int M0 = HALF ? 4 : 0, M1 = HALF ? 12 : 16;
    __shared__ float2 buf[PL];
    constexpr int TWO = fft::last_stage_tw_offset<L>();
    __shared__ float2 twm[TWO > 0 ? TWO : 1];
    const int t = threadIdx.x;
    for (int i = t; i < TWO; i += TPL) twm[i] = tw[i];
    float2 wl[16];
#pragma unroll
    for (int r = 1; r < 16; ++r) wl[r] = tw[TWO + (r - 1) * TPL + t];
    // the field is at most L x L complex64 (128 MiB for L = 4096)
    const __amdgpu_buffer_rsrc_t src = oph_buf_rsrc(in, (unsigned)((size_t)nx * ny * sizeof(float2)));
    float2 nxt[M1 - M0];
    int r = blockIdx.x;
    auto fetch = [&](int row) {
#pragma unroll
        for (int m = M0; m < M1; ++m) {
            const int c = t + m * TPL - ox;
            const bool ok = row < ny && c >= 0 && c < nx;
            nxt[m - M0] = oph_buf_load_f2(src, ok ? (unsigned)(((size_t)row * nx + c) * sizeof(float2)) : OPH_BUF_OOB);
        }
    };
    auto row = [&](int r) {
        float2 v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = (m >= M0 && m < M1) ? nxt[m - M0] : make_float2(0.f, 0.f);
        __syncthreads();  // the previous row's last-stage reads are done
        fft::stage0_from_regs<L, -1>(v, buf, t);
        fetch(r + gridDim.x);  // into the registers this row was read from
        __syncthreads();
        fft::middle_stages_tw<L, -1>(buf, twm, t);
        fft::last_stage_to_regs_w<L, -1>(buf, wl, t, v);
        float2* dst = work + (size_t)(oy + r) * L;
#pragma unroll
        for (int m = 0; m < 16; ++m) dst[t + m * TPL] = v[m];
    };
    fetch(r);
    if (r < ny) {
        row(r);
        for (r += gridDim.x; r < ny; r += gridDim.x) row(r);
    }
}

template <int L, bool HALF>
__global__ __launch_bounds__(L / 16) void asm_rows_inv_f(const float2* __restrict__ work, int nx, int ny, int ox,
                                                         int oy, float scale, float2* __restrict__ out,
                                                         const float2* __restrict__ tw) {
    constexpr int TPL = L / 16;
    constexpr int PL = fft::padded_len(L);
    constexpr int M0 = HALF ? 4 : 0, M1 = HALF ? 12 : 16;
    __shared__ float2 buf[PL];
    constexpr int TWO = fft::last_stage_tw_offset<L>();
    __shared__ float2 twm[TWO > 0 ? TWO : 1];
    const int t = threadIdx.x;
    for (int i = t; i < TWO; i += TPL) twm[i] = tw[i];
    // last-stage twiddles W^(r t) rebuilt per row from W^t and W^(4t) (two
    // registers pairs instead of 15: four waves per SIMD)
    const float2 w1 = tw[TWO + t], w4 = tw[TWO + 3 * TPL + t];
    // rows oy .. oy + ny - 1 of the work grid and the output field, each at
    // most L x L complex64 (128 MiB for L = 4096)
    const __amdgpu_buffer_rsrc_t src =
        oph_buf_rsrc(work + (size_t)oy * L, (unsigned)((size_t)ny * L * sizeof(float2)));
    const __amdgpu_buffer_rsrc_t dst = oph_buf_rsrc(out, (unsigned)((size_t)nx * ny * sizeof(float2)));
    float2 nxt[16];
    int r = blockIdx.x;
    auto fetch = [&](int row) {
#pragma unroll
        for (int m = 0; m < 16; ++m)
            nxt[m] = oph_buf_load_f2(
                src, row < ny ? (unsigned)(((size_t)row * L + t + m * TPL) * sizeof(float2)) : OPH_BUF_OOB);
    };
    auto row = [&](int r) {
        float2 v[16];
#pragma unroll
        for (int m = 0; m < 16; ++m) v[m] = nxt[m];
        __syncthreads();
        fft::stage0_from_regs<L, +1>(v, buf, t);
        fetch(r + gridDim.x);  // into the registers this row was read from
        __syncthreads();
        fft::middle_stages_tw<L, +1>(buf, twm, t);
        float2 wl[16];
        fft::twiddle_powers(w1, w4, wl);
        fft::last_stage_to_regs_w<L, +1>(buf, wl, t, v);
#pragma unroll
        for (int m = M0; m < M1; ++m) {
            const int c = t + m * TPL - ox;
            const bool ok = c >= 0 && c < nx;
            oph_buf_store_f2(make_float2(v[m].x * scale, v[m].y * scale), dst,
                             ok ? (unsigned)(((size_t)r * nx + c) * sizeof(float2)) : OPH_BUF_OOB);
        }
    };
    fetch(r);
    if (r < ny) {
        row(r);
        for (r += gridDim.x; r < ny; r += gridDim.x) row(r);
    }
}

// resident workgroups per CU of the persistent row passes (register-bound
// occupancy of each kernel, 4 waves per workgroup)
constexpr int ROWS_FWD_PER_CU = 2, ROWS_INV_PER_CU = 3;
// workgroups per CU of the register-fused row passes: at most 3, measured
// faster than the occupancy maximum of 4 (2048 rows of 4096: forward 17.4
// against 19.0 us, inverse 18.9 against 19.2 us; 2 per CU: 19.0 and 20.1 us)
constexpr int ROWS_F_MAX_PER_CU = 3;

// Workgroups of a persistent kernel resident per CU, from the runtime's
// occupancy calculator (registers and LDS of the compiled kernel).
template <typename K>
int resident_per_cu(K kernel, int threads) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void*>(kernel), threads, 0) !=
            hipSuccess ||
        n < 1) {
        hipGetLastError();
        n = 1;
    }
    return n;
}

template <int LX>
int launch_rows_fwd(oph_ctx* ctx, const float2* in, int nx, int ny, int ox, int oy, float2* work, const TabJob& tj) {
    const float2* tw = oph_twiddles(ctx, LX);
    if (!tw) return OPH_ENOMEM;
    constexpr int LPW = fftk::row_lines(LX);
    if constexpr (LPW == 1) {
        const int g = std::min(ny, ROWS_FWD_PER_CU * std::max(ctx->cu_count, 1));
        if constexpr (fft::regs_fusable<LX>()) {
            static const int occ_h = std::min(ROWS_F_MAX_PER_CU, resident_per_cu(asm_rows_fwd_f<LX, true>, LX / 16));
            static const int occ_f = std::min(ROWS_F_MAX_PER_CU, resident_per_cu(asm_rows_fwd_f<LX, false>, LX / 16));
            const int cus = std::max(ctx->cu_count, 1);
            if (ox >= LX / 4 && ox + nx <= 3 * LX / 4)
                OPH_LAUNCH(ctx, "asm_rows_fwd", (asm_rows_fwd_f<LX, true>), dim3(std::min(ny, occ_h * cus)),
                           dim3(LX / 16), 0, in, nx, ny, ox, oy, work, tw, tj);
            else
                OPH_LAUNCH(ctx, "asm_rows_fwd", (asm_rows_fwd_f<LX, false>), dim3(std::min(ny, occ_f * cus)),
                           dim3(LX / 16), 0, in, nx, ny, ox, oy, work, tw, tj);
            return OPH_OK;
        }
        OPH_LAUNCH(ctx, "asm_rows_fwd", asm_rows_fwd_p<LX>, dim3(g), dim3(LX / 16), 0, in, nx, ny, ox, oy, work, tw,
                   tj);
        return OPH_OK;
    }
    OPH_LAUNCH(ctx, "asm_rows_fwd", asm_rows_fwd<LX>, dim3((ny + LPW - 1) / LPW), dim3(LPW * (LX / 16)), 0, in,
               nx, ny, ox, oy, work, tw, tj);
    return OPH_OK;
}

template <int LX>
int launch_rows_inv(oph_ctx* ctx, const float2* work, int nx, int ny, int ox, int oy, float scale, float2* out) {
    const float2* tw = oph_twiddles(ctx, LX);
    if (!tw) return OPH_ENOMEM;
    constexpr int LPW = fftk::row_lines(LX);
    if constexpr (LPW == 1) {
        const int g = std::min(ny, ROWS_INV_PER_CU * std::max(ctx->cu_count, 1));
        if constexpr (fft::regs_fusable<LX>()) {
            static const int occ_h = std::min(ROWS_F_MAX_PER_CU, resident_per_cu(asm_rows_inv_f<LX, true>, LX / 16));
            static const int occ_f = std::min(ROWS_F_MAX_PER_CU, resident_per_cu(asm_rows_inv_f<LX, false>, LX / 16));
            const int cus = std::max(ctx->cu_count, 1);
            if (ox >= LX / 4 && ox + nx <= 3 * LX / 4)
                OPH_LAUNCH(ctx, "asm_rows_inv", (asm_rows_inv_f<LX, true>), dim3(std::min(ny, occ_h * cus)),
                           dim3(LX / 16), 0, work, nx, ny, ox, oy, scale, out, tw);
            else
                OPH_LAUNCH(ctx, "asm_rows_inv", (asm_rows_inv_f<LX, false>), dim3(std::min(ny, occ_f * cus)),
                           dim3(LX / 16), 0, work, nx, ny, ox, oy, scale, out, tw);
            return OPH_OK;
        }
        OPH_LAUNCH(ctx, "asm_rows_inv", asm_rows_inv_p<LX>, dim3(g), dim3(LX / 16), 0, work, nx, ny, ox, oy, scale,
                   out, tw);
        return OPH_OK;
    }
    OPH_LAUNCH(ctx, "asm_rows_inv", asm_rows_inv<LX>, dim3((ny + LPW - 1) / LPW), dim3(LPW * (LX / 16)), 0, work,
               nx, ny, ox, oy, scale, out, tw);
    return OPH_OK;
}

// (the band limit is a separate instantiation: its row-table terms would
// otherwise take registers in the unlimited kernel too)
template <int B, bool EXACT, bool BAND>
int launch_cols_mid_t(oph_ctx* ctx, float2* work, int npx, const float4* tx, const float4* ty, const AsmConst& k,
                      const float2* tw) {
    constexpr int NTHR = B * (MID_L / 16);
    const size_t shm = ((size_t)B * fft::padded_len(MID_L) + MID_TW) * sizeof(float2);
    static const int occ = [&] {
        hipFuncSetAttribute(reinterpret_cast<const void*>(&asm_cols_mid<B, EXACT, BAND>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        hipGetLastError();
        int n = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, reinterpret_cast<const void*>(&asm_cols_mid<B, EXACT, BAND>), NTHR, shm) != hipSuccess ||
            n < 1) {
            hipGetLastError();
            n = 1;
        }
        return n;
    }();
    // one resident wave of workgroups, a multiple of 8, at most one per group
    const int grid = std::max(8, std::min(npx / B, occ * std::max(ctx->cu_count, 1)) / 8 * 8);
    const float2* ty2 = reinterpret_cast<const float2*>(ty + MID_L);
    OPH_LAUNCH(ctx, "asm_cols", (asm_cols_mid<B, EXACT, BAND>), dim3(grid), dim3(NTHR), shm, work, npx, tx, ty, ty2,
               k, tw);
    return OPH_OK;
}

template <int B, bool EXACT>
int launch_cols_mid(oph_ctx* ctx, float2* work, int npx, const float4* tx, const float4* ty, const AsmConst& k,
                    const float2* tw) {
    return k.band ? launch_cols_mid_t<B, EXACT, true>(ctx, work, npx, tx, ty, k, tw)
                  : launch_cols_mid_t<B, EXACT, false>(ctx, work, npx, tx, ty, k, tw);
}

template <int LY, int B, bool EXACT>
int launch_cols_b(oph_ctx* ctx, float2* work, int npx, int ny, int oy, const float4* tx, const float4* ty,
                  const AsmConst& k, const float2* tw) {
    const size_t shm = ((size_t)B * fft::padded_len(LY) + (asm_cols_fused<LY>() ? 256 : 0)) * sizeof(float2);
    static bool attr = false;
    if (!attr) {
        hipFuncSetAttribute(reinterpret_cast<const void*>(&asm_cols<LY, B, EXACT>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        hipGetLastError();
        attr = true;
    }
    // the compact y table follows the two float4 tables (ty has LY entries)
    const float2* ty2 = reinterpret_cast<const float2*>(ty + LY);
    OPH_LAUNCH(ctx, "asm_cols", (asm_cols<LY, B, EXACT>), dim3((npx + B - 1) / B), dim3(B * (LY / 16)), shm,
               work, npx, ny, oy, tx, ty, ty2, k, tw);
    return OPH_OK;
}

template <int LY, bool EXACT>
int launch_cols_ex(oph_ctx* ctx, float2* work, int npx, int ny, int oy, const float4* tx, const float4* ty,
                   const AsmConst& k) {
    const float2* tw = oph_twiddles(ctx, LY);
    if (!tw) return OPH_ENOMEM;
    constexpr int BMAX = fftk::col_lines(LY);
    int b = fftk::col_width(LY, BMAX);
    switch (b) {
        case 1: return launch_cols_b<LY, 1, EXACT>(ctx, work, npx, ny, oy, tx, ty, k, tw);
        case 2:
            if constexpr (BMAX >= 2) {
                if constexpr (LY == MID_L) {
                    if (oy == LY / 4 && ny == LY / 2 && npx % (8 * MID_B) == 0)
                        return launch_cols_mid<MID_B, EXACT>(ctx, work, npx, tx, ty, k, tw);
                }
                return launch_cols_b<LY, 2, EXACT>(ctx, work, npx, ny, oy, tx, ty, k, tw);
            }
            break;
        case 4:
            if constexpr (BMAX >= 4) return launch_cols_b<LY, 4, EXACT>(ctx, work, npx, ny, oy, tx, ty, k, tw);
            break;
        case 8:
            if constexpr (BMAX >= 8) return launch_cols_b<LY, 8, EXACT>(ctx, work, npx, ny, oy, tx, ty, k, tw);
            break;
        default:
            if constexpr (BMAX >= 16) return launch_cols_b<LY, 16, EXACT>(ctx, work, npx, ny, oy, tx, ty, k, tw);
            break;
    }
    return launch_cols_b<LY, BMAX, EXACT>(ctx, work, npx, ny, oy, tx, ty, k, tw);
}

template <int LY>
int launch_cols(oph_ctx* ctx, bool exact, float2* work, int npx, int ny, int oy, const float4* tx,
                const float4* ty, const AsmConst& k) {
    return exact ? launch_cols_ex<LY, true>(ctx, work, npx, ny, oy, tx, ty, k)
                 : launch_cols_ex<LY, false>(ctx, work, npx, ny, oy, tx, ty, k);
}

}  // namespace

// Device-pointer core of oph_asm_propagate: field (ny x nx) in place, FFT grid
// npx x npy (powers of two >= nx, ny), work buffer of npx*npy.
int oph_asm_device(oph_ctx* ctx, float2* field, int nx, int ny, int npx, int npy, double px, double py,
                   double lambda, double z, int band_limit, float2* work) {
    const int ox = npx / 2 - nx / 2, oy = npy / 2 - ny / 2;
    // x and y tables (float4), then the y table again as float2
    float4* tab = static_cast<float4*>(
        oph_scratch(ctx, SLOT_TABLE, sizeof(float4) * (npx + npy) + sizeof(float2) * npy));
    if (!tab) return OPH_ENOMEM;
    const TabJob tj{tab, reinterpret_cast<float2*>(tab + npx + npy), npx, npy, px, py, lambda, z};
    const double zl = z / lambda;
    AsmConst k;
    k.zl = zl;
    k.c0 = (float)(zl - rint(zl));
    k.K = (float)zl;
    k.dfx = 1.0 / (npx * px);
    k.dfy = 1.0 / (npy * py);
    k.lam2 = lambda * lambda;
    k.npx = npx;
    k.npy = npy;
    k.band = band_limit;
    // truncation of the series at u^5: next term 21/1024 u^6 * z/lambda
    const double umax = lambda * lambda * (0.25 / (px * px) + 0.25 / (py * py));
    const bool exact = umax >= 0.5 || fabs(zl) * pow(umax, 6) * (21.0 / 1024.0) > 1e-7;
    int rc;
    switch (npx) {
#define OPH_RF(n) case n: rc = launch_rows_fwd<n>(ctx, field, nx, ny, ox, oy, work, tj); break;
        OPH_FFT_LENGTHS(OPH_RF)
#undef OPH_RF
        default: return oph_fail(ctx, OPH_ENOTSUP, "ASM grid width %d unsupported", npx);
    }
    if (rc) return rc;
    switch (npy) {
#define OPH_CF(n) case n: rc = launch_cols<n>(ctx, exact, work, npx, ny, oy, tab, tab + npx, k); break;
        OPH_FFT_LENGTHS(OPH_CF)
#undef OPH_CF
        default: return oph_fail(ctx, OPH_ENOTSUP, "ASM grid height %d unsupported", npy);
    }
    if (rc) return rc;
    const float scale = (float)(1.0 / ((double)npx * npy));
    switch (npx) {
#define OPH_RI(n) case n: rc = launch_rows_inv<n>(ctx, work, nx, ny, ox, oy, scale, field); break;
        OPH_FFT_LENGTHS(OPH_RI)
#undef OPH_RI
    }
    return rc;
}

extern "C" int oph_asm_propagate(oph_ctx* ctx, oph_c64* field, int nx, int ny, double px, double py,
                                 double lambda, double z, int pad, int band_limit, int flags) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!field) return oph_fail(ctx, OPH_EINVAL, "null field");
    if (pad != 1 && pad != 2) return oph_fail(ctx, OPH_EINVAL, "pad must be 1 or 2");
    if (!(px > 0) || !(py > 0) || !(lambda > 0)) return oph_fail(ctx, OPH_EINVAL, "bad pitch or wavelength");
    const int npx = nx * pad, npy = ny * pad;
    if (!fftk::supported(npx) || !fftk::supported(npy))
        return oph_fail(ctx, OPH_ENOTSUP, "padded grid %dx%d: sides must be powers of two in 16..8192", npx, npy);
    if ((flags & OPH_ASYNC) && !(flags & OPH_DEVICE_PTR))
        return oph_fail(ctx, OPH_EINVAL, "OPH_ASYNC needs device pointers");
    const size_t fbytes = (size_t)nx * ny * sizeof(float2);
    float2* f = reinterpret_cast<float2*>(field);
    if (!(flags & OPH_DEVICE_PTR)) {
        f = static_cast<float2*>(oph_scratch(ctx, SLOT_IN, fbytes));
        if (!f) return OPH_ENOMEM;
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(f, field, fbytes, hipMemcpyHostToDevice, ctx->stream));
    }
    float2* work = static_cast<float2*>(oph_scratch(ctx, SLOT_WORK1, (size_t)npx * npy * sizeof(float2)));
    if (!work) return OPH_ENOMEM;
    if ((rc = oph_asm_device(ctx, f, nx, ny, npx, npy, px, py, lambda, z, band_limit, work))) return rc;
    if (!(flags & OPH_DEVICE_PTR))
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(field, f, fbytes, hipMemcpyDeviceToHost, ctx->stream));
    return oph_finish(ctx, flags);
}
