// cnn.hip -- the learned correction of a coarse depth-map hologram: a small
// residual CNN (3x3 convolutions) with its forward and backward passes and an
// Adam step, all as HIP kernels on gfx950.  There is no counterpart in the
// reference; the network corrects a 16-layer depth-map field toward the
// 256-layer one (both produced by oph_depthmap).
//
// Network (channels C, blocks NB; real/imaginary parts as 2 channels):
//   h0 = conv_head(x)                       2 -> C
//   h_b = h_{b-1} + conv_b2(relu(conv_b1(h_{b-1})))     b = 1..NB
//   y  = x + conv_tail(h_NB)                C -> 2
// Tensors are NHWC float32 in HBM.  A convolution is an implicit GEMM on
// MFMA: M = pixels, N = output channels, K = 9 taps x input channels; each
// wave produces 32 pixels x 32 channels with v_mfma_f32_32x32x16_bf16,
// operands rounded to BF16 (weights and activations), FP32 accumulation;
// bias, ReLU and the residual add are fused into the epilogue.  The input
// gradient is the same convolution with flipped taps and transposed
// channels; the weight gradient is a GEMM over pixels (M = output channels,
// N = 9 x input channels + 1 bias column) split into pixel chunks whose
// partial sums are reduced in a fixed order (deterministic).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <type_traits>
#include <vector>

#include "oph_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));


struct oph_cnn {
    int blocks = 8, ch = 32;
    int64_t nparams = 0;
    float* params = nullptr;  // device, fp32
    float* grads = nullptr;
    float* m = nullptr;       // Adam moments
    float* v = nullptr;
    bf16x8* wf = nullptr;     // BF16 B fragments of every forward convolution
    bf16x8* wb = nullptr;     // ... of every input-gradient convolution (flipped taps, transposed channels)
    int64_t step = 0;
    // parameter revision and the revision the forward fragments were packed
    // from (inference skips the packing while they agree)
    uint64_t params_rev = 1, wf_rev = 0;
    // activation workspace
    float* act = nullptr;
    size_t act_bytes = 0;
    float* part = nullptr;    // weight-gradient partials
    size_t part_bytes = 0;
    double* red = nullptr;    // loss reduction
};

namespace {

constexpr int CMAX = 32;

// parameter offsets: head W[C][9][2], b[C]; block b: W1[C][9][C], b1[C],
// W2[C][9][C], b2[C]; tail W[2][9][C], b[2]
struct Layout {
    int C, NB;
    __host__ __device__ int64_t head_w() const { return 0; }
    __host__ __device__ int64_t head_b() const { return (int64_t)C * 9 * 2; }
    __host__ __device__ int64_t block_base(int b) const {
        return head_b() + C + (int64_t)b * (2 * ((int64_t)C * 9 * C + C));
    }
    __host__ __device__ int64_t w1(int b) const { return block_base(b); }
    __host__ __device__ int64_t b1(int b) const { return w1(b) + (int64_t)C * 9 * C; }
    __host__ __device__ int64_t w2(int b) const { return b1(b) + C; }
    __host__ __device__ int64_t b2(int b) const { return w2(b) + (int64_t)C * 9 * C; }
    __host__ __device__ int64_t tail_w() const { return block_base(NB); }
    __host__ __device__ int64_t tail_b() const { return tail_w() + 2 * 9 * (int64_t)C; }
    __host__ __device__ int64_t total() const { return tail_b() + 2; }
};

// ---- forward convolution (implicit GEMM) -------------------------------------

// TI / TO: element type of the input / output (and residual) tensors,
// float or __bf16 (inference keeps the activations between layers in BF16:
// the MFMA operands are BF16-rounded either way)
template <typename TI = float, typename TO = float>
struct ConvArgsT {
    const TI* in;         // [B][H][W][CIN]
    const bf16x8* wf;     // B fragments [KSTEPS][64 lanes] (pack_weights)
    const float* bias;    // [COUT] or null
    const TO* res;        // residual added after the bias (or null), [..][COUT]
    TO* out;              // [B][H][W][COUT]
    int B, H, W;
    int relu;
};
typedef ConvArgsT<float, float> ConvArgs;

// A operand for lane (pixel p, k0..k0+7): taps/channels of the input patch
template <int CIN>
__device__ __forceinline__ bf16x8 load_patch(const ConvArgs& a, int b, int y, int x, int k0) {
    bf16x8 v;
    if constexpr (CIN % 8 == 0) {
        const int tap = k0 / CIN, c = k0 % CIN;
        const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
        if (tap < 9 && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) {
            const float4* s =
                reinterpret_cast<const float4*>(a.in + (((size_t)b * a.H + yy) * a.W + xx) * CIN + c);
            const float4 p0 = s[0], p1 = s[1];
            v[0] = (__bf16)p0.x; v[1] = (__bf16)p0.y; v[2] = (__bf16)p0.z; v[3] = (__bf16)p0.w;
            v[4] = (__bf16)p1.x; v[5] = (__bf16)p1.y; v[6] = (__bf16)p1.z; v[7] = (__bf16)p1.w;
        } else {
            v = (bf16x8)(__bf16)0.f;
        }
    } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = k0 + e;
            const int tap = k / CIN, c = k % CIN;
            const int yy = y + tap / 3 - 1, xx = x + tap % 3 - 1;
            float f = 0.f;
            if (tap < 9 && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
                f = a.in[(((size_t)b * a.H + yy) * a.W + xx) * CIN + c];
            v[e] = (__bf16)f;
        }
    }
    return v;
}

// Tiled forward convolution.  A workgroup of 8 waves owns an 8-row x 32-
// column pixel tile of one image: it stages the tile and its one-pixel halo
// in LDS once, as BF16 (pixel stride padded by 16 bytes so that the 16-byte
// A-fragment reads of a half-wave fall on distinct banks), and each wave
// keeps its B fragments (the weights, BF16) in registers for the whole tile.
// Wave w computes tile rows 2w and 2w + 1: 32 pixels x 32 output channels
// each, K = 9 CIN in steps of 16 on v_mfma_f32_32x32x16_bf16 (one weight
// fragment feeds two MFMAs).  Lane l: pixel l%32, k half l/32;
// accumulator register v of lane l is pixel 8(v/4) + 4(l/32) + v%4 of the
// row, channel l%32.
constexpr int CT_H = 8, CT_W = 32;
constexpr int CT_WAVES = CT_H / 2;

__host__ __device__ constexpr int ksteps_c(int cin) { return (9 * cin + 15) / 16; }

template <int CIN>
__host__ __device__ constexpr int tile_stride() { return CIN % 8 == 0 ? CIN + 8 : CIN; }

template <int CIN, int COUT, typename TI = float, typename TO = float>
__global__ __launch_bounds__(64 * CT_WAVES) void conv3x3_fwd(ConvArgsT<TI, TO> a) {
    constexpr int K = 9 * CIN;
    constexpr int KSTEPS = (K + 15) / 16;
    constexpr int PS = tile_stride<CIN>();
    constexpr int TWH = CT_W + 2;
    __shared__ __attribute__((aligned(16))) __bf16 tile[(CT_H + 2) * TWH * PS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x0 = blockIdx.x * CT_W, y0 = blockIdx.y * CT_H, b = blockIdx.z;
    const size_t img = (size_t)b * a.H * a.W;

    // stage the tile (+ halo) as BF16; outside the image is zero padding
    if constexpr (std::is_same<TI, __bf16>::value && CIN % 8 == 0) {
        constexpr int Q = CIN / 8;  // 16-byte vectors per pixel
        for (int i = threadIdx.x; i < (CT_H + 2) * TWH * Q; i += 64 * CT_WAVES) {
            const int pix = i / Q, q = i % Q;
            const int yy = y0 + pix / TWH - 1, xx = x0 + pix % TWH - 1;
            bf16x8 v = (bf16x8)(__bf16)0.f;
            if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
                v = reinterpret_cast<const bf16x8*>(a.in + (img + (size_t)yy * a.W + xx) * CIN)[q];
            *reinterpret_cast<bf16x8*>(tile + pix * PS + 8 * q) = v;
        }
    } else if constexpr (CIN % 4 == 0 && std::is_same<TI, float>::value) {
        constexpr int Q = CIN / 4;  // float4 per pixel
        for (int i = threadIdx.x; i < (CT_H + 2) * TWH * Q; i += 64 * CT_WAVES) {
            const int pix = i / Q, q = i % Q;
            const int yy = y0 + pix / TWH - 1, xx = x0 + pix % TWH - 1;
            float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
                v = reinterpret_cast<const float4*>(a.in + (img + (size_t)yy * a.W + xx) * CIN)[q];
            bf16x4 w4;
            w4[0] = (__bf16)v.x;
            w4[1] = (__bf16)v.y;
            w4[2] = (__bf16)v.z;
            w4[3] = (__bf16)v.w;
            *reinterpret_cast<bf16x4*>(tile + pix * PS + 4 * q) = w4;
        }
    } else {
        for (int i = threadIdx.x; i < (CT_H + 2) * TWH * CIN; i += 64 * CT_WAVES) {
            const int pix = i / CIN, c = i % CIN;
            const int yy = y0 + pix / TWH - 1, xx = x0 + pix % TWH - 1;
            float v = 0.f;
            if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) v = (float)a.in[(img + (size_t)yy * a.W + xx) * CIN + c];
            tile[pix * PS + c] = (__bf16)v;
        }
    }
    // weights -> registers (one 16-byte fragment per k step)
    const int n = lane & 31, kh = lane >> 5;
    bf16x8 Bf[KSTEPS];
#pragma unroll
    for (int s = 0; s < KSTEPS; ++s) Bf[s] = a.wf[s * 64 + lane];
    __syncthreads();

    const int col = lane & 31;
    f32x16 acc[2] = {(f32x16)0.f, (f32x16)0.f};
#pragma unroll
    for (int s = 0; s < KSTEPS; ++s) {
        const int k0 = 16 * s + 8 * kh;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int r = 2 * wave + h;
            bf16x8 A;
            if constexpr (CIN % 16 == 0) {
                // k = 16 s + 8 kh stays within tap 16 s / CIN: compile-time
                // tap and channel base, the lane's half added at run time
                const int tap = (16 * s) / CIN, c = (16 * s) % CIN + 8 * kh;
                const __bf16* src = tile + ((r + tap / 3) * TWH + col + tap % 3) * PS + c;
                A = *reinterpret_cast<const bf16x8*>(src);
            } else if constexpr (CIN % 8 == 0) {
                const int tap = k0 / CIN, c = k0 % CIN;
                if (tap < 9) {
                    const __bf16* src = tile + ((r + tap / 3) * TWH + col + tap % 3) * PS + c;
                    A = *reinterpret_cast<const bf16x8*>(src);
                } else {
                    A = (bf16x8)(__bf16)0.f;
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int k = k0 + e, tap = k / CIN, c = k % CIN;
                    A[e] = tap < 9 ? tile[((r + tap / 3) * TWH + col + tap % 3) * PS + c] : (__bf16)0.f;
                }
            }
            acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Bf[s], A, acc[h], 0, 0, 0);
        }
    }
    // epilogue: the weights are the MFMA A operand, so register v of lane
    // (col, kh) is output channel 8 (v/4) + 4 kh + v%4 of pixel col: four
    // runs of four consecutive channels per lane (vector stores, one bounds
    // check per pixel)
    constexpr int NV = COUT < 4 ? COUT : 4;  // channels per run (COUT is 2 or a multiple of 4)
    static_assert(COUT < 4 || COUT % 4 == 0, "output channels");
    const int x = x0 + col;
    if (x >= a.W) return;
    // residuals first, all in flight together (out may not be assumed
    // distinct from res by the compiler, which would otherwise serialise
    // each load behind the previous store)
    float rv[2][4][4];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int ch0 = 8 * q + 4 * kh, y = y0 + 2 * wave + h;
#pragma unroll
            for (int i = 0; i < 4; ++i) rv[h][q][i] = 0.f;
            if (!a.res || y >= a.H || ch0 >= COUT) continue;
            const TO* r = a.res + (img + (size_t)y * a.W + x) * COUT + ch0;
#pragma unroll
            for (int i = 0; i < NV; ++i) rv[h][q][i] = (float)r[i];
        }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int ch0 = 8 * q + 4 * kh;
        if (ch0 >= COUT) continue;
        float bias[4] = {0.f, 0.f, 0.f, 0.f};
        if (a.bias) {
#pragma unroll
            for (int i = 0; i < NV; ++i) bias[i] = a.bias[ch0 + i];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int y = y0 + 2 * wave + h;
            if (y >= a.H) continue;
            TO o[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = acc[h][4 * q + i] + bias[i];
                if (a.relu) v = fmaxf(v, 0.f);
                o[i] = (TO)(v + rv[h][q][i]);
            }
            TO* dst = a.out + (img + (size_t)y * a.W + x) * COUT + ch0;
            if constexpr (NV == 4) {
                if constexpr (std::is_same<TO, float>::value)
                    *reinterpret_cast<float4*>(dst) = make_float4(o[0], o[1], o[2], o[3]);
                else
                    *reinterpret_cast<bf16x4*>(dst) = bf16x4{o[0], o[1], o[2], o[3]};
            } else {
#pragma unroll
                for (int i = 0; i < NV; ++i) dst[i] = o[i];
            }
        }
    }
}

// B fragments of convolution layers: fragment (s, lane) holds, for output
// channel n = lane % 32 and k = 16 s + 8 (lane / 32) + e (e = 0..7; k = tap
// CIN + c), W[n][tap][c] rounded to BF16 (0 beyond COUT or 9 CIN).  For an
// input-gradient convolution (flip) the weights of the original layer
// (CIN_o = cout, COUT_o = cin) are read with flipped taps and transposed
// channels: W'[n][tap][c] = W_o[c][8 - tap][n].  One launch packs every layer
// of a pass (blockIdx.y = layer).
struct PackLayer {
    int64_t woff, foff;   // parameter offset of W_o, fragment offset
    int cin, cout, flip;  // channels of the convolution as applied
};
constexpr int MAX_BLOCKS = 16;
constexpr int MAX_PACK = 2 * MAX_BLOCKS + 2;
struct PackPlan {
    PackLayer l[MAX_PACK];
};

__global__ void pack_weights(const float* __restrict__ P, bf16x8* __restrict__ out, PackPlan plan) {
    const PackLayer L = plan.l[blockIdx.y];
    const int K = 9 * L.cin, ks = (K + 15) / 16;
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= ks * 64) return;
    const int s = f / 64, lane = f % 64, n = lane & 31, k0 = 16 * s + 8 * (lane >> 5);
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int k = k0 + e;
        float w = 0.f;
        if (n < L.cout && k < K) {
            const int tap = k / L.cin, c = k % L.cin;
            w = L.flip ? P[L.woff + ((int64_t)c * 9 + (8 - tap)) * L.cout + n]
                       : P[L.woff + ((int64_t)n * 9 + tap) * L.cin + c];
        }
        v[e] = (__bf16)w;
    }
    out[L.foff + f] = v;
}

// relu'(pre) from the post-activation: g *= [act > 0]
__global__ void relu_mask(float* g, const float* act, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(act[i] > 0.f)) g[i] = 0.f;
}

// ---- weight gradient -------------------------------------------------------------
// dW[n][k] = sum_p dY[p][n] X[p + off(tap k)][c k], db[n] = sum_p dY[p][n].
// GEMM M = COUT (padded to 32), N = 9 CIN + 1 (bias column), K = pixels.
// Pixels are taken in 8-row x 32-column tiles of one image.  A workgroup
// stages a tile in LDS as BF16: dY transposed ([channel][pixel], so the 8
// pixels of an A fragment are 16 contiguous bytes) and X as three copies of
// its channel planes shifted by dx = -1, 0, +1 (so the 8 pixels of a B
// fragment at any tap are 16 aligned bytes).  Five waves take two of the ten
// 32-column N tiles each.  Workgroup g accumulates tiles g, g + G, ... in
// registers (a fixed order) and writes part[g][32][NCOL]; dw_reduce sums the
// partials in workgroup order (deterministic).

constexpr int DW_WAVES = 5;
constexpr int DW_MAX_GROUPS = 256;

struct DwArgs {
    const float* dy;   // [npix][COUT]
    const float* x;    // [npix][CIN]
    float* part;       // [groups][32][ncol]
    int B, H, W;
    int tiles_x, tiles_y;
};

template <int CIN, int COUT>
__global__ __launch_bounds__(64 * DW_WAVES) void conv3x3_dw(DwArgs a) {
    constexpr int NCOL = 9 * CIN + 1;
    constexpr int NT = (NCOL + 31) / 32;
    constexpr int TPW = (NT + DW_WAVES - 1) / DW_WAVES;  // N tiles per wave
    constexpr int TP = CT_H * CT_W;                      // pixels per tile
    constexpr int XR = CT_H + 2;                         // X rows incl. halo
    __shared__ __attribute__((aligned(16))) __bf16 sdy[32 * TP];
    __shared__ __attribute__((aligned(16))) __bf16 sx[3 * CIN * XR * CT_W];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 31, kh = 8 * (lane >> 5);
    f32x16 acc[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] = (f32x16)0.f;
    const int ntiles = a.tiles_x * a.tiles_y * a.B;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int tx = tile % a.tiles_x, ty = (tile / a.tiles_x) % a.tiles_y, b = tile / (a.tiles_x * a.tiles_y);
        const int x0 = tx * CT_W, y0 = ty * CT_H;
        const size_t img = (size_t)b * a.H * a.W;
        __syncthreads();  // the previous tile's reads are done
        for (int i = threadIdx.x; i < TP * 32; i += 64 * DW_WAVES) {
            const int px = i / 32, n = i % 32;
            const int yy = y0 + px / CT_W, xx = x0 + px % CT_W;
            float v = 0.f;
            if (n < COUT && yy < a.H && xx < a.W) v = a.dy[(img + (size_t)yy * a.W + xx) * COUT + n];
            sdy[n * TP + px] = (__bf16)v;
        }
        for (int i = threadIdx.x; i < XR * (CT_W + 2) * CIN; i += 64 * DW_WAVES) {
            const int px = i / CIN, c = i % CIN;
            const int r = px / (CT_W + 2), j = px % (CT_W + 2);   // j = 0..33 <-> x0 - 1 + j
            const int yy = y0 + r - 1, xx = x0 + j - 1;
            float v = 0.f;
            if (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) v = a.x[(img + (size_t)yy * a.W + xx) * CIN + c];
            const __bf16 bv = (__bf16)v;
            // shifted copy d (dx = d - 1) holds column jj = j - d at [d][c][r][jj]
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const int jj = j - d;
                if (jj >= 0 && jj < CT_W) sx[((d * CIN + c) * XR + r) * CT_W + jj] = bv;
            }
        }
        __syncthreads();
#pragma unroll
        for (int tt = 0; tt < TPW; ++tt) {
            const int t = wave + DW_WAVES * tt;
            if (t >= NT) break;
            const int col = 32 * t + row;   // B column of this lane: (tap, c) or the bias column
            const int tap = col / CIN, c = col % CIN;
            const bool is_w = col < NCOL - 1, is_b = col == NCOL - 1;
            const int dyo = tap / 3, dxo = tap % 3;  // row offset into the halo rows, shifted copy
#pragma unroll 4
            for (int s = 0; s < TP / 16; ++s) {
                const int p0 = 16 * s + kh;               // 8 pixels of one tile row
                const int pr = p0 / CT_W, pc = p0 % CT_W;
                const bf16x8 A = *reinterpret_cast<const bf16x8*>(sdy + row * TP + p0);
                bf16x8 Bv;
                if (is_w)
                    Bv = *reinterpret_cast<const bf16x8*>(sx + ((dxo * CIN + c) * XR + pr + dyo) * CT_W + pc);
                else
                    Bv = (bf16x8)(__bf16)(is_b ? 1.f : 0.f);
                acc[tt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A, Bv, acc[tt], 0, 0, 0);
            }
        }
    }
    float* dst = a.part + (size_t)blockIdx.x * 32 * NCOL;
#pragma unroll
    for (int tt = 0; tt < TPW; ++tt) {
        const int t = wave + DW_WAVES * tt;
        if (t >= NT) break;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int r = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);  // output channel
            const int col = 32 * t + (lane & 31);
            if (col < NCOL) dst[(size_t)r * NCOL + col] = acc[tt][v];
        }
    }
}

// Sum of the workgroup partials into the gradient (W and b): one wave per
// element; lane j adds partials j, j + 64, ... in order, then the lanes are
// combined by a fixed butterfly (the same order on every run).
__global__ __launch_bounds__(256) void dw_reduce(const float* part, int groups, int cout, int ncol, float* gw,
                                                 float* gb) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= cout * ncol) return;
    const int r = i / ncol, col = i % ncol;
    double s = 0.0;
    for (int g = lane; g < groups; g += 64) s += part[((size_t)g * 32 + r) * ncol + col];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane != 0) return;
    if (col == ncol - 1)
        gb[r] = (float)s;
    else
        gw[(size_t)r * (ncol - 1) + col] = (float)s;
}

// ---- loss, Adam ------------------------------------------------------------------

// dy = 2 (y - t) / n, per-block partial sums of (y - t)^2 (fixed order)
__global__ void mse_grad(const float* y, const float* t, float* dy, int64_t n, double* part) {
    __shared__ double s[256];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double e2 = 0.0;
    if (i < n) {
        const float d = y[i] - t[i];
        dy[i] = 2.f * d / (float)n;
        e2 = (double)d * d;
    }
    s[threadIdx.x] = e2;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void sum_partials(double* part, int n) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += part[i];
        part[n] = s;
    }
}

__global__ void adam_step(float* p, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2,
                          float eps, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void add_inplace(float* a, const float* b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] += b[i];
}

// ---- fused residual block (inference, BF16 activations) ----------------------
// h' = h + conv2(relu(conv1(h))) for C = 32 channels in one kernel.  A
// persistent workgroup of 4 waves (conv1's weight fragments in registers,
// conv2's in LDS, loaded once) walks 8 x 32 output tiles; for each it
// stages h on the tile plus a two-pixel halo (12 x 36 pixels) in LDS,
// computes relu(conv1) on the tile plus a one-pixel halo (10 x 34 pixels:
// ten rows of 32 and a 10 x 2 strip; pixels outside the image are the zero
// padding of conv2) into a second LDS tile, then conv2 on the 8 rows of 32
// with the residual read from the staged h.  Each wave takes consecutive
// rows so that one LDS operand serves the taps of up to three output rows.
// HBM traffic per block: h read once (plus halo), h' written once.  The
// MFMAs take the weights as the A
// operand, so each lane ends with 4 runs of 4 consecutive channels of one
// pixel (8-byte stores, one bounds check per pixel).  Same fragments,
// BF16 rounding and FP32 accumulation as conv3x3_fwd.
constexpr int RB_TH = 8, RB_TW = 32;
constexpr int RB_IH = RB_TH + 4, RB_IW = RB_TW + 4;   // staged input
constexpr int RB_MH = RB_TH + 2, RB_MW = RB_TW + 2;   // conv1 output
constexpr int RB_MID = RB_MH * RB_MW;                 // 340 pixels
constexpr int RB_PS = 32 + 8;                         // LDS pixel stride (bf16)
constexpr int RB_NPRE = (RB_IH * RB_IW * 4 + 255) / 256;  // 16-byte pieces per thread (7)

struct ResArgs {
    const __bf16* h;      // [B][H][W][32] (unused by the head form)
    __bf16* out;          // [B][H][W][32], distinct from h
    const bf16x8* w1;     // B fragments of conv1 / conv2 (pack_weights)
    const bf16x8* w2;
    const float* b1;
    const float* b2;
    int B, H, W;
    // head form: h = conv_head(x) computed on the staged tile
    const float* x;       // [B][H][W][2]
    const bf16x8* w0;     // head fragments (2 k steps)
    const float* b0;
};

// Head form of the first block: instead of h, the 2-channel network input x
// is staged (14 x 38 pixels, a three-pixel halo, as BF16 pairs) and the head
// convolution computes h on the 12 x 36 staged region, in 14 groups of 32
// pixels (K = 18 in two k steps of 16: k = 2 tap + channel, as conv3x3_fwd
// orders it), BF16-rounded exactly as the separate head layer stores it and
// zero outside the image.  h never goes to HBM.
constexpr int RB_XH = RB_IH + 2, RB_XW = RB_IW + 2;    // staged input of the head form
constexpr int RB_NPX = (RB_XH * RB_XW + 255) / 256;    // its 8-byte pieces per thread (3)
constexpr int RB_HG = (RB_IH * RB_IW + 31) / 32;       // head groups of 32 pixels (14)

// Software pipelines of the two convolutions, unrolled at compile time: the
// LDS operands of step S are read at step S and consumed by the MFMAs at
// step S + AHEAD, and sched_group_barrier keeps that order (masks: 0x100 LDS
// read, 0x008 MFMA).
constexpr int RB_C1_AHEAD = 6, RB_C2_AHEAD = 3;

// conv1 steps: S = (NR + 2) k + i takes operand k = 2 dx + half of input
// row rb_c1_row<NR>(i); the row order makes consecutive MFMAs alternate
// between the output rows' accumulators (no back-to-back dependent MFMAs):
// NR = 3: rows 2, 0, 3, 1, 4; NR = 2: rows 0, 3, 1, 2.
template <int NR>
__host__ __device__ constexpr int rb_c1_row(int i) {
    return NR == 3 ? (i == 0 ? 2 : i == 1 ? 0 : i == 2 ? 3 : i == 3 ? 1 : 4) : (i == 0 ? 0 : i == 1 ? 3 : i == 2 ? 1 : 2);
}
template <int NR>
__host__ __device__ constexpr int rb_c1_mfmas(int u) {
    const int ri = rb_c1_row<NR>(u % (NR + 2));
    int n = 0;
    for (int j = 0; j < NR; ++j) n += (ri - j >= 0 && ri - j <= 2) ? 1 : 0;
    return n;
}

template <int NR, int S>
__device__ __forceinline__ void rb_c1_pipe(const __bf16* src, bf16x8 (&X)[RB_C1_AHEAD + 1], f32x16 (&acc)[NR],
                                           const bf16x8 (&Bf)[18]) {
    constexpr int NS = (NR + 2) * 6, AH = RB_C1_AHEAD;
    if constexpr (S < NS + AH) {
        if constexpr (S < NS) {
            constexpr int ri = rb_c1_row<NR>(S % (NR + 2)), k = S / (NR + 2), dx = k >> 1, half = k & 1;
            X[S % (AH + 1)] = *reinterpret_cast<const bf16x8*>(src + (ri * RB_IW + dx) * RB_PS + 16 * half);
        }
        constexpr int U = S - AH;
        if constexpr (U >= 0) {
            constexpr int ri = rb_c1_row<NR>(U % (NR + 2)), k = U / (NR + 2), dx = k >> 1, half = k & 1;
#pragma unroll
            for (int j = 0; j < NR; ++j) {
                const int dy = ri - j;
                if (dy >= 0 && dy <= 2)
                    acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Bf[(3 * dy + dx) * 2 + half], X[U % (AH + 1)],
                                                                     acc[j], 0, 0, 0);
            }
        }
        if constexpr (S < NS) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        if constexpr (U >= 0) __builtin_amdgcn_sched_group_barrier(0x008, rb_c1_mfmas<NR>(U), 0);
        rb_c1_pipe<NR, S + 1>(src, X, acc, Bf);
    }
}

// conv2 step S = 6 dy + k: weight fragment S, operands of mid rows r0 + dy
// (read at step S - 6, or here for dy = 0) and r0 + dy + 1 (read here)
// `late` issues the epilogue's NL LDS reads once the operand reads are done,
// so that they complete under the last steps' MFMAs.
template <int S, int NL, typename F>
__device__ __forceinline__ void rb_c2_pipe(const __bf16* src, const bf16x8* w2l, bf16x8 (&Xr)[4][6],
                                           bf16x8 (&Wr)[RB_C2_AHEAD + 1], f32x16 (&acc)[2], F& late) {
    constexpr int NS = 18, AH = RB_C2_AHEAD;
    if constexpr (S < NS + AH) {
        if constexpr (S == NS && NL > 0) {
            late();
            __builtin_amdgcn_sched_group_barrier(0x100, NL, 0);
        }
        if constexpr (S < NS) {
            constexpr int dy = S / 6, k = S % 6, dx = k >> 1, half = k & 1;
            Wr[S % (AH + 1)] = w2l[S * 64];
            if constexpr (dy == 0) Xr[0][k] = *reinterpret_cast<const bf16x8*>(src + dx * RB_PS + 16 * half);
            Xr[dy + 1][k] = *reinterpret_cast<const bf16x8*>(src + ((dy + 1) * RB_MW + dx) * RB_PS + 16 * half);
        }
        constexpr int U = S - AH;
        if constexpr (U >= 0) {
            constexpr int dy = U / 6, k = U % 6;
            acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Wr[U % (AH + 1)], Xr[dy][k], acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Wr[U % (AH + 1)], Xr[dy + 1][k], acc[1], 0, 0, 0);
        }
        if constexpr (S < NS) __builtin_amdgcn_sched_group_barrier(0x100, S < 6 ? 3 : 2, 0);
        if constexpr (U >= 0) __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
        rb_c2_pipe<S + 1, NL>(src, w2l, Xr, Wr, acc, late);
    }
}

// relu(acc + bias) of one lane's 16 channels (4 runs of 4) as BF16, all
// zero when keep is 0 (pixels outside the image: conv2's zero padding).
// Two values per conversion; the relu is taken on the BF16 pair as a signed
// 16-bit maximum with 0 (a negative BF16, -0 included, has the sign bit
// set), which equals rounding relu(x) and is one instruction per pair; a
// mask instead of a branch; the bias is already in registers, so nothing
// here waits on LDS.
__device__ __forceinline__ unsigned rb_relu_bf16x2(f32x2 v) {
    const i16x2 h = __builtin_bit_cast(i16x2, __builtin_convertvector(v, bf16x2));
    return __builtin_bit_cast(unsigned, __builtin_elementwise_max(h, (i16x2)0));
}
__device__ __forceinline__ void rb_relu_store(const f32x16& acc, const float4 (&bq)[4], unsigned keep,
                                              __bf16* dst) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const f32x2 lo = {acc[4 * q + 0] + bq[q].x, acc[4 * q + 1] + bq[q].y};
        const f32x2 hi = {acc[4 * q + 2] + bq[q].z, acc[4 * q + 3] + bq[q].w};
        u32x2 o;
        o[0] = rb_relu_bf16x2(lo) & keep;
        o[1] = rb_relu_bf16x2(hi) & keep;
        *reinterpret_cast<u32x2*>(dst + 8 * q) = o;
    }
}

// conv1 (weights in registers) + bias + relu for NR consecutive rows of 32
// pixels of the halo tile, mid rows r0 .. r0 + NR - 1, mid columns px.
// Input row r0 + ri feeds output row r0 + j at tap row dy = ri - j, so
// each 16-byte operand (input row, column offset dx, channel half) is read
// from LDS once and used by up to 3 MFMAs.
template <int NR>
__device__ __forceinline__ void rb_conv1_rows(const __bf16* in_t, __bf16* mid_t, const bf16x8 (&Bf)[18],
                                              const float4* bias_s, int r0, int px, int kh, int x0, int y0, int H,
                                              int W) {
    float4 bq[4];  // bias of this lane's channels 8 q + 4 kh + (0..3)
#pragma unroll
    for (int q = 0; q < 4; ++q) bq[q] = bias_s[2 * q + kh];
    __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);  // their own group, ahead of the pipeline's
    f32x16 acc[NR];
#pragma unroll
    for (int j = 0; j < NR; ++j) acc[j] = (f32x16)0.f;
    const __bf16* src = in_t + (r0 * RB_IW + px) * RB_PS + 8 * kh;
    bf16x8 X[RB_C1_AHEAD + 1];
    rb_c1_pipe<NR, 0>(src, X, acc, Bf);
    const int xx = x0 - 1 + px;
    const bool xin = xx >= 0 && xx < W;
#pragma unroll
    for (int j = 0; j < NR; ++j) {
        const int mr = r0 + j;
        const int yy = y0 - 1 + mr;
        const unsigned keep = xin && yy >= 0 && yy < H ? ~0u : 0u;  // else conv2's zero padding
        rb_relu_store(acc[j], bq, keep, mid_t + (mr * RB_MW + px) * RB_PS + 4 * kh);
    }
}

// The strip's 18 k steps as a software pipeline like rb_c1_pipe: the operand
// of step S is read at step S and used by the MFMA at step S + AHEAD.
template <int S>
__device__ __forceinline__ void rb_strip_pipe(const __bf16* src, bf16x8 (&X)[RB_C1_AHEAD + 1], f32x16 (&ac2)[2],
                                              const bf16x8 (&Bf)[18]) {
    constexpr int NS = 18, AH = RB_C1_AHEAD;
    if constexpr (S < NS + AH) {
        if constexpr (S < NS) {
            constexpr int tap = S >> 1;
            X[S % (AH + 1)] =
                *reinterpret_cast<const bf16x8*>(src + ((tap / 3) * RB_IW + tap % 3) * RB_PS + 16 * (S & 1));
        }
        constexpr int U = S - AH;
        if constexpr (U >= 0)
            ac2[U & 1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Bf[U], X[U % (AH + 1)], ac2[U & 1], 0, 0, 0);
        if constexpr (S < NS) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        if constexpr (U >= 0) __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        rb_strip_pipe<S + 1>(src, X, ac2, Bf);
    }
}

// conv1 for the 10 x 2 strip of mid columns 32, 33 (20 pixels; lanes 20..31
// repeat the last one and store nothing).
__device__ __forceinline__ void rb_conv1_strip(const __bf16* in_t, __bf16* mid_t, const bf16x8 (&Bf)[18],
                                               const float4* bias_s, int px, int kh, int x0, int y0, int H, int W) {
    const int f = px < 2 * RB_MH ? px : 2 * RB_MH - 1;
    const int mr = f >> 1, mc = RB_TW + (f & 1);
    float4 bq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) bq[q] = bias_s[2 * q + kh];
    __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);  // their own group, ahead of the pipeline's
    const __bf16* src = in_t + (mr * RB_IW + mc) * RB_PS + 8 * kh;
    // two accumulation chains (even and odd k steps) so that consecutive
    // MFMAs do not depend on each other
    f32x16 ac2[2] = {(f32x16)0.f, (f32x16)0.f};
    bf16x8 X[RB_C1_AHEAD + 1];
    rb_strip_pipe<0>(src, X, ac2, Bf);
    const f32x16 acc = ac2[0] + ac2[1];
    if (px >= 2 * RB_MH) return;
    const int yy = y0 - 1 + mr, xx = x0 - 1 + mc;
    const unsigned keep = yy >= 0 && yy < H && xx >= 0 && xx < W ? ~0u : 0u;
    rb_relu_store(acc, bq, keep, mid_t + (mr * RB_MW + mc) * RB_PS + 4 * kh);
}

// Outputs of one wave for one tile (2 rows x 4 runs of 4 channels, BF16)
// and their byte offsets in the image (out of range for pixels outside it).
struct RbOut {
    u32x2 v[2][4];
    unsigned dst[2];
};

__device__ __forceinline__ void rb_store(const RbOut& o, __amdgpu_buffer_rsrc_t img) {
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) __builtin_amdgcn_raw_buffer_store_b64(o.v[j][q], img, o.dst[j] + 16 * q, 0, 0);
}

// conv2 (weights from LDS) + bias + residual for output rows r0, r0 + 1 of
// the tile.  Taps are taken row by row (dy outer): the operands of mid row
// r0 + dy + 1 serve output row r0 + 1 at dy and output row r0 at dy + 1, so
// each is read once; each weight fragment is read once for both rows.
__device__ __forceinline__ void rb_conv2_rows(const __bf16* in_t, const __bf16* mid_t, const bf16x8* w2_s,
                                              const float4* bias_s, RbOut& out, int r0, int lane, int x0, int y0,
                                              int H, int W) {
    const int px = lane & 31, kh = lane >> 5;
    f32x16 acc[2] = {(f32x16)0.f, (f32x16)0.f};
    const __bf16* src = mid_t + (r0 * RB_MW + px) * RB_PS + 8 * kh;
    bf16x8 Xr[4][6], Wr[RB_C2_AHEAD + 1];
    // bias and residual, read under the last MFMAs
    float4 bq[4];
    bf16x4 rv[2][4];
    const __bf16* res = in_t + ((r0 + 2) * RB_IW + px + 2) * RB_PS + 4 * kh;
    auto late = [&]() {
#pragma unroll
        for (int q = 0; q < 4; ++q) bq[q] = bias_s[8 + 2 * q + kh];
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) rv[j][q] = *reinterpret_cast<const bf16x4*>(res + j * RB_IW * RB_PS + 8 * q);
    };
    rb_c2_pipe<0, 12>(src, w2_s + lane, Xr, Wr, acc, late);
    const int x = x0 + px;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int r = r0 + j;
        const int y = y0 + r;
        // pixels outside the image: offset out of range, store dropped
        out.dst[j] = y < H && x < W ? (unsigned)((y * W + x) * 64 + 8 * kh) : OPH_BUF_OOB;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            bf16x4 o;
            o[0] = (__bf16)(acc[j][4 * q + 0] + bq[q].x + (float)rv[j][q][0]);
            o[1] = (__bf16)(acc[j][4 * q + 1] + bq[q].y + (float)rv[j][q][1]);
            o[2] = (__bf16)(acc[j][4 * q + 2] + bq[q].z + (float)rv[j][q][2]);
            o[3] = (__bf16)(acc[j][4 * q + 3] + bq[q].w + (float)rv[j][q][3]);
            out.v[j][q] = __builtin_bit_cast(u32x2, o);
        }
    }
}

// Head convolution of the head form: groups g = wave, wave + 4, ... of 32
// staged pixels p = 32 g + px (row p / 36, column p % 36 of the 12 x 36
// region); lane half kh holds k = 16 s + 8 kh + e, i.e. taps 8 s + 4 kh ..
// + 3 (two channels each, one 4-byte read per tap; taps beyond 8 are zero).
__device__ __forceinline__ void rb_head(const unsigned* x_t, __bf16* in_t, const bf16x8 (&B0)[2],
                                        const float4* bias0, int wave, int px, int kh, int x0, int y0, int H,
                                        int W) {
    float4 bq[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) bq[q] = bias0[2 * q + kh];
#pragma unroll
    for (int gi = 0; gi < (RB_HG + 3) / 4; ++gi) {
        const int g = wave + 4 * gi;
        if (g >= RB_HG) break;  // uniform per wave
        const int p = 32 * g + px;
        const int pc = p < RB_IH * RB_IW ? p : RB_IH * RB_IW - 1;
        const int r = pc / RB_IW, c = pc % RB_IW;
        const unsigned* src = x_t + r * RB_XW + c;
        u32x4 v0, v1 = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int tap = 4 * kh + e;  // k step 0: taps 0..7
            v0[e] = src[(tap / 3) * RB_XW + tap % 3];
        }
        if (kh == 0) v1[0] = src[2 * RB_XW + 2];  // k step 1: tap 8 (lane half 0 only)
        f32x16 acc = (f32x16)0.f;
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(B0[0], __builtin_bit_cast(bf16x8, v0), acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(B0[1], __builtin_bit_cast(bf16x8, v1), acc, 0, 0, 0);
        const int yy = y0 - 2 + r, xx = x0 - 2 + c;
        const unsigned keep = p < RB_IH * RB_IW && yy >= 0 && yy < H && xx >= 0 && xx < W ? ~0u : 0u;
        __bf16* dst = in_t + p * RB_PS + 4 * kh;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            // bias, then + 0 as the separate layer adds its (absent) residual
            const f32x2 lo = {(acc[4 * q + 0] + bq[q].x) + 0.f, (acc[4 * q + 1] + bq[q].y) + 0.f};
            const f32x2 hi = {(acc[4 * q + 2] + bq[q].z) + 0.f, (acc[4 * q + 3] + bq[q].w) + 0.f};
            u32x2 o;
            o[0] = __builtin_bit_cast(unsigned, __builtin_convertvector(lo, bf16x2)) & keep;
            o[1] = __builtin_bit_cast(unsigned, __builtin_convertvector(hi, bf16x2)) & keep;
            if (p < RB_IH * RB_IW) *reinterpret_cast<u32x2*>(dst + 8 * q) = o;
        }
    }
}

template <bool HEAD>
__global__ __launch_bounds__(256, 2) void resblock_fwd(ResArgs a) {
    constexpr int KSTEPS = ksteps_c(32);
    // staged input plus room for the unused tail of the last 16-byte pieces
    // (the staging writes are unconditional)
    __shared__ __attribute__((aligned(16))) __bf16 in_t[(RB_NPRE * 256 / 4) * RB_PS];
    __shared__ __attribute__((aligned(16))) __bf16 mid_t[RB_MID * RB_PS];
    __shared__ bf16x8 w2_s[KSTEPS * 64];
    __shared__ float4 bias_s[24];  // b1, b2, b0 (head form)
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: scalar branches
    // conv1's fragments in registers, conv2's in LDS, loaded once for every
    // tile this (persistent) workgroup takes
    bf16x8 Bf[KSTEPS];
#pragma unroll
    for (int s = 0; s < KSTEPS; ++s) Bf[s] = a.w1[s * 64 + lane];
    for (int i = threadIdx.x; i < KSTEPS * 64; i += 256) w2_s[i] = a.w2[i];
    if (threadIdx.x < (HEAD ? 24 : 16))
        bias_s[threadIdx.x] =
            reinterpret_cast<const float4*>(threadIdx.x < 8 ? a.b1 : threadIdx.x < 16 ? a.b2 : a.b0)[threadIdx.x & 7];
    bf16x8 B0[2];  // head fragments (head form)
    if constexpr (HEAD) {
        B0[0] = a.w0[lane];
        B0[1] = a.w0[64 + lane];
    }
    const int tiles_x = (a.W + RB_TW - 1) / RB_TW, tiles_y = (a.H + RB_TH - 1) / RB_TH;
    const int ntiles = tiles_x * tiles_y * a.B;
    // The staged input of the next tile (12 x 36 pixels x 64 bytes, 16-byte
    // pieces i = threadIdx.x + 256 n) is loaded into registers while the
    // current tile computes; the epilogues read only LDS, so no wait on
    // global loads falls inside the convolutions.
    constexpr int NPRE = HEAD ? 0 : RB_NPRE;
    constexpr int NPX = HEAD ? RB_NPX : 0;
    bf16x8 pre[NPRE > 0 ? NPRE : 1];
    u32x2 prex[NPX > 0 ? NPX : 1];
    const unsigned img_bytes = (unsigned)a.H * (unsigned)a.W * 64u;  // < 2^31 (checked by the launcher)
    auto fetch = [&](int tile) {
        const int tx = tile % tiles_x, ty = (tile / tiles_x) % tiles_y, b = tile / (tiles_x * tiles_y);
        const int x0 = tx * RB_TW, y0 = ty * RB_TH;
        if constexpr (HEAD) {
            const __amdgpu_buffer_rsrc_t rx = oph_buf_rsrc(a.x + (size_t)b * a.H * a.W * 2, img_bytes / 8);
#pragma unroll
            for (int n = 0; n < NPX; ++n) {
                const int i = threadIdx.x + 256 * n;
                const int yy = y0 + i / RB_XW - 3, xx = x0 + i % RB_XW - 3;
                const bool ok = i < RB_XH * RB_XW && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
                const unsigned off = ok ? (unsigned)((yy * a.W + xx) * 8) : OPH_BUF_OOB;
                prex[n] = __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(rx, off, 0, 0));
            }
            return;
        }
        const __amdgpu_buffer_rsrc_t rs = oph_buf_rsrc(a.h + (size_t)b * a.H * a.W * 32, img_bytes);
#pragma unroll
        for (int n = 0; n < NPRE; ++n) {
            const int i = threadIdx.x + 256 * n;
            const int pix = i >> 2, q = i & 3;
            const int yy = y0 + pix / RB_IW - 2, xx = x0 + pix % RB_IW - 2;
            const bool ok = i < RB_IH * RB_IW * 4 && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
            const unsigned off = ok ? (unsigned)((yy * a.W + xx) * 64 + 16 * q) : OPH_BUF_OOB;
            pre[n] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
        }
    };
    // A tile's outputs are stored after the next tile's staging and before
    // its prefetch, so that the prefetch is always the newest memory
    // operation when the staging waits for it (the wait counter is in order)
    // and the stores have a whole tile to complete.
    RbOut pend;
    __amdgpu_buffer_rsrc_t pend_img = oph_buf_rsrc(a.out, img_bytes);
    if ((int)blockIdx.x < ntiles) fetch(blockIdx.x);
    // the grid stride kept in a register (otherwise it is reloaded from the
    // dispatch packet every tile, with a wait that also drains LDS)
    int gstep = (int)gridDim.x;
    asm volatile("" : "+s"(gstep));
    for (int tile = blockIdx.x; tile < ntiles; tile += gstep) {
        const int tx = tile % tiles_x, ty = (tile / tiles_x) % tiles_y, b = tile / (tiles_x * tiles_y);
        const int x0 = tx * RB_TW, y0 = ty * RB_TH;
        const size_t img = (size_t)b * a.H * a.W;
        // lane and wave indices through opaque copies: everything derived
        // from them is recomputed per tile instead of being hoisted out of
        // the loop into registers held for the kernel's lifetime
        int lane_o = lane, wave_o = wave;
        asm volatile("" : "+v"(lane_o), "+s"(wave_o));
        const int px = lane_o & 31, kh = lane_o >> 5;
        __syncthreads();  // the previous tile's LDS reads are done
        // head form: the staged input (BF16 pairs) in the mid tile's space,
        // which is free until conv1
        unsigned* x_t = reinterpret_cast<unsigned*>(mid_t);
        if constexpr (HEAD) {
#pragma unroll
            for (int n = 0; n < NPX; ++n) {
                const int i = threadIdx.x + 256 * n;
                const f32x2 v = __builtin_bit_cast(f32x2, prex[n]);
                if (i < RB_XH * RB_XW) x_t[i] = __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
            }
        } else {
#pragma unroll
            for (int n = 0; n < NPRE; ++n) {
                const int i = threadIdx.x + 256 * n;
                *reinterpret_cast<bf16x8*>(in_t + (i >> 2) * RB_PS + 8 * (i & 3)) = pre[n];
            }
        }
        __syncthreads();
        if (tile != (int)blockIdx.x) rb_store(pend, pend_img);
        if (tile + gstep < ntiles) fetch(tile + gstep);
        if constexpr (HEAD) {
            rb_head(x_t, in_t, B0, bias_s + 16, wave_o, px, kh, x0, y0, a.H, a.W);
            __syncthreads();
        }

        // conv1 + bias + relu on the 10 x 34 halo tile: rows of 32 (mid
        // columns 0..31) as 3, 3, 2, 2 consecutive rows per wave, and the
        // 10 x 2 strip of columns 32, 33 as one flat group on wave 3
        if (wave_o < 2)
            rb_conv1_rows<3>(in_t, mid_t, Bf, bias_s, 3 * wave_o, px, kh, x0, y0, a.H, a.W);
        else
            rb_conv1_rows<2>(in_t, mid_t, Bf, bias_s, 6 + 2 * (wave_o - 2), px, kh, x0, y0, a.H, a.W);
        if (wave_o == 3) rb_conv1_strip(in_t, mid_t, Bf, bias_s, px, kh, x0, y0, a.H, a.W);
        __syncthreads();

        // conv2 + bias + residual on the 8 x 32 tile: rows 2 wave, 2 wave + 1
        rb_conv2_rows(in_t, mid_t, w2_s, bias_s, pend, 2 * wave_o, lane_o, x0, y0, a.H, a.W);
        pend_img = oph_buf_rsrc(a.out + img * 32, img_bytes);
    }
    if ((int)blockIdx.x < ntiles) rb_store(pend, pend_img);
}

// ---- pipelined residual block: two wave groups ----------------------------------
// The same block as resblock_fwd, one workgroup of 8 waves per CU with the
// work split by phase instead of by tile: waves 0-3 (group A) compute
// relu(conv1) of tile i + 1 into one mid buffer and stage tile i + 2's input
// (prefetched into registers one tile earlier) into one input buffer, while
// waves 4-7 (group B) compute conv2 of tile i from the other mid buffer and
// store it, its residual read into registers one tile earlier; one barrier
// per tile.  Each group keeps only its own convolution's weight fragments,
// so both sets live in registers and no weight passes through LDS (in
// resblock_fwd conv2's fragments are read from LDS for every tile, a fifth
// of the block's LDS reads).  LDS: two staged inputs and two mid tiles
// (123 KB).  The arithmetic (fragments, k order, rounding) is
// resblock_fwd's.
constexpr int RP_IN = (RB_NPRE * 256 / 4) * RB_PS;  // bf16 per staged-input buffer
constexpr int RP_MID = RB_MID * RB_PS;              // bf16 per mid buffer

// Tiles of a persistent workgroup g of gstep: with gstep a multiple of 8,
// each XCD (workgroups are placed on the 8 XCDs round-robin by index) walks
// its own contiguous share of the tiles in row-major order with its
// workgroups side by side, so that the halo columns neighbouring tiles share
// and the halo rows shared with the previous tile row come from its own L2;
// otherwise tiles g, g + gstep, ...  rb_count: how many; rb_tile: the k-th.
struct RbSched {
    int g, gstep, per, x, s, share, lim;
    __device__ RbSched(int g_, int gstep_, int ntiles) : g(g_), gstep(gstep_) {
        per = gstep % 8 == 0 ? gstep / 8 : 0;
        x = g % 8;
        s = g / 8;
        share = (ntiles + 7) / 8;
        lim = per ? min(share, ntiles - x * share) : ntiles;
    }
    __device__ int count() const {
        if (!per) return g < lim ? (lim - 1 - g) / gstep + 1 : 0;
        return lim > s ? (lim - 1 - s) / per + 1 : 0;
    }
    __device__ int tile(int k) const { return per ? x * share + s + k * per : g + k * gstep; }
};

__device__ __forceinline__ void rb_read_res(const __bf16* in_t, int r0, int px, int kh, bf16x4 (&rv)[2][4]) {
    const __bf16* res = in_t + ((r0 + 2) * RB_IW + px + 2) * RB_PS + 4 * kh;
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) rv[j][q] = *reinterpret_cast<const bf16x4*>(res + j * RB_IW * RB_PS + 8 * q);
}

// conv2 pipeline with the weight fragments in registers (rb_c2_pipe's
// operand order and MFMA order)
template <int S>
__device__ __forceinline__ void rb_c2r_pipe(const __bf16* src, const bf16x8 (&W)[18], bf16x8 (&Xr)[4][6],
                                            f32x16 (&acc)[2]) {
    constexpr int NS = 18, AH = 6;
    if constexpr (S < NS + AH) {
        if constexpr (S < NS) {
            constexpr int dy = S / 6, k = S % 6, dx = k >> 1, half = k & 1;
            if constexpr (dy == 0) Xr[0][k] = *reinterpret_cast<const bf16x8*>(src + dx * RB_PS + 16 * half);
            Xr[dy + 1][k] = *reinterpret_cast<const bf16x8*>(src + ((dy + 1) * RB_MW + dx) * RB_PS + 16 * half);
        }
        constexpr int U = S - AH;
        if constexpr (U >= 0) {
            constexpr int dy = U / 6, k = U % 6;
            acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W[U], Xr[dy][k], acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(W[U], Xr[dy + 1][k], acc[1], 0, 0, 0);
        }
        if constexpr (S < NS) __builtin_amdgcn_sched_group_barrier(0x100, S < 6 ? 2 : 1, 0);
        if constexpr (U >= 0) __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
        rb_c2r_pipe<S + 1>(src, W, Xr, acc);
    }
}

// conv2 + bias + residual (from registers) for output rows r0, r0 + 1
__device__ __forceinline__ void rb_conv2_regs(const __bf16* mid_t, const bf16x8 (&W)[18], const float4 (&bq)[4],
                                              const bf16x4 (&rv)[2][4], RbOut& out, int r0, int lane, int x0,
                                              int y0, int H, int W_) {
    const int px = lane & 31, kh = lane >> 5;
    f32x16 acc[2] = {(f32x16)0.f, (f32x16)0.f};
    const __bf16* src = mid_t + (r0 * RB_MW + px) * RB_PS + 8 * kh;
    bf16x8 Xr[4][6];
    rb_c2r_pipe<0>(src, W, Xr, acc);
    const int x = x0 + px;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int y = y0 + r0 + j;
        out.dst[j] = y < H && x < W_ ? (unsigned)((y * W_ + x) * 64 + 8 * kh) : OPH_BUF_OOB;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            bf16x4 o;
            o[0] = (__bf16)(acc[j][4 * q + 0] + bq[q].x + (float)rv[j][q][0]);
            o[1] = (__bf16)(acc[j][4 * q + 1] + bq[q].y + (float)rv[j][q][1]);
            o[2] = (__bf16)(acc[j][4 * q + 2] + bq[q].z + (float)rv[j][q][2]);
            o[3] = (__bf16)(acc[j][4 * q + 3] + bq[q].w + (float)rv[j][q][3]);
            out.v[j][q] = __builtin_bit_cast(u32x2, o);
        }
    }
}

__global__ __launch_bounds__(512, 1) void resblock_pipe(ResArgs a) {
    constexpr int KSTEPS = ksteps_c(32);
    __shared__ __attribute__((aligned(16))) __bf16 in_t[2 * RP_IN];
    __shared__ __attribute__((aligned(16))) __bf16 mid_t[2 * RP_MID];
    __shared__ float4 bias_s[8];  // b1 (conv1's epilogue reads it from LDS)
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (threadIdx.x < 8) bias_s[threadIdx.x] = reinterpret_cast<const float4*>(a.b1)[threadIdx.x];
    const int tiles_x = (a.W + RB_TW - 1) / RB_TW, tiles_y = (a.H + RB_TH - 1) / RB_TH;
    const int ntiles = tiles_x * tiles_y * a.B;
    int gstep = (int)gridDim.x;
    asm volatile("" : "+s"(gstep));
    const RbSched sched(blockIdx.x, gstep, ntiles);
    const int nmy = sched.count();
    const unsigned img_bytes = (unsigned)a.H * (unsigned)a.W * 64u;  // < 2^31 (checked by the launcher)
    auto tile_of = [&](int k, int& x0, int& y0, int& b) {
        const int tile = sched.tile(k);
        x0 = (tile % tiles_x) * RB_TW;
        y0 = ((tile / tiles_x) % tiles_y) * RB_TH;
        b = tile / (tiles_x * tiles_y);
    };
    if (wave < 4) {
        // group A: conv1 and the staging (threads 0..255, as in resblock_fwd)
        bf16x8 Bf[KSTEPS];
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) Bf[s] = a.w1[s * 64 + lane];
        // two tiles' input in flight: tile t is prefetched into pa (t even)
        // or pb (t odd) two tiles before it is staged (one tile ahead, the
        // staging waited on memory latency: the block without any MFMA
        // took 37 of its ~60 us)
        bf16x8 pa[RB_NPRE], pb[RB_NPRE];
        auto fetch = [&](int k, bf16x8 (&pre)[RB_NPRE]) {
            int x0, y0, b;
            tile_of(k, x0, y0, b);
            const __amdgpu_buffer_rsrc_t rs = oph_buf_rsrc(a.h + (size_t)b * a.H * a.W * 32, img_bytes);
#pragma unroll
            for (int n = 0; n < RB_NPRE; ++n) {
                const int i = threadIdx.x + 256 * n;
                const int pix = i >> 2, q = i & 3;
                const int yy = y0 + pix / RB_IW - 2, xx = x0 + pix % RB_IW - 2;
                const bool ok = i < RB_IH * RB_IW * 4 && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
                const unsigned off = ok ? (unsigned)((yy * a.W + xx) * 64 + 16 * q) : OPH_BUF_OOB;
                pre[n] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
            }
        };
        auto stage = [&](int buf, const bf16x8 (&pre)[RB_NPRE]) {
#pragma unroll
            for (int n = 0; n < RB_NPRE; ++n) {
                const int i = threadIdx.x + 256 * n;
                *reinterpret_cast<bf16x8*>(in_t + buf * RP_IN + (i >> 2) * RB_PS + 8 * (i & 3)) = pre[n];
            }
        };
        auto conv1 = [&](int k) {
            int x0, y0, b;
            tile_of(k, x0, y0, b);
            int lane_o = lane, wave_o = wave;
            asm volatile("" : "+v"(lane_o), "+s"(wave_o));
            const int px = lane_o & 31, kh = lane_o >> 5;
            const __bf16* it = in_t + (k & 1) * RP_IN;
            __bf16* mt = mid_t + (k & 1) * RP_MID;
            if (wave_o < 2)
                rb_conv1_rows<3>(it, mt, Bf, bias_s, 3 * wave_o, px, kh, x0, y0, a.H, a.W);
            else
                rb_conv1_rows<2>(it, mt, Bf, bias_s, 6 + 2 * (wave_o - 2), px, kh, x0, y0, a.H, a.W);
            if (wave_o == 3) rb_conv1_strip(it, mt, Bf, bias_s, px, kh, x0, y0, a.H, a.W);
        };
        if (nmy > 0) {
            fetch(0, pa);
            stage(0, pa);
        }
        if (nmy > 1) fetch(1, pb);
        if (nmy > 2) fetch(2, pa);
        __syncthreads();
        if (nmy > 0) conv1(0);
        if (nmy > 1) stage(1, pb);
        if (nmy > 3) fetch(3, pb);
        __syncthreads();
        // two tiles per trip (the prefetch arrays alternate); one barrier
        // per tile, nmy in all, as group B
        for (int k = 0; k < nmy; k += 2) {
            if (k + 1 < nmy) conv1(k + 1);
            if (k + 2 < nmy) stage(0, pa);  // tile k + 2 into the buffer tile k used
            if (k + 4 < nmy) fetch(k + 4, pa);
            __syncthreads();
            if (k + 1 >= nmy) break;
            if (k + 2 < nmy) conv1(k + 2);
            if (k + 3 < nmy) stage(1, pb);
            if (k + 5 < nmy) fetch(k + 5, pb);
            __syncthreads();
        }
    } else {
        // group B: conv2 (fragments and bias in registers), the residual
        // one tile ahead, the stores
        bf16x8 W2[KSTEPS];
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) W2[s] = a.w2[s * 64 + lane];
        float4 bq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) bq[q] = reinterpret_cast<const float4*>(a.b2)[2 * q + (lane >> 5)];
        bf16x4 rv[2][4];
        const int r0 = 2 * (wave - 4);
        __syncthreads();
        if (nmy > 0) rb_read_res(in_t, r0, lane & 31, lane >> 5, rv);
        __syncthreads();
        for (int k = 0; k < nmy; ++k) {
            int x0, y0, b;
            tile_of(k, x0, y0, b);
            int lane_o = lane;
            asm volatile("" : "+v"(lane_o));
            RbOut out;
            rb_conv2_regs(mid_t + (k & 1) * RP_MID, W2, bq, rv, out, r0, lane_o, x0, y0, a.H, a.W);
            rb_store(out, oph_buf_rsrc(a.out + (size_t)b * a.H * a.W * 32, img_bytes));
            if (k + 1 < nmy) rb_read_res(in_t + ((k + 1) & 1) * RP_IN, r0, lane_o & 31, lane_o >> 5, rv);
            __syncthreads();
        }
    }
}

// ---- tail convolution (inference, BF16 input) ----------------------------------
// y = x + conv_tail(h), 32 -> 2 channels.  A persistent workgroup stages h
// on 8 x 32 tiles with a one-pixel halo (10 x 34 pixels: the geometry of the
// residual block's mid tile) through prefetch registers and runs the block's
// conv2 pipeline on it with the tail's fragments in LDS; lanes of the first
// half hold output channels 0 and 1.  Same fragments, MFMA order and
// rounding as conv3x3_fwd<32, 2> (whose results it reproduces bit for bit),
// but the staging of the next tile overlaps the current one's MFMAs.
constexpr int TL_NPRE = (RB_MH * RB_MW * 4 + 255) / 256;  // 16-byte pieces per thread (6)

struct TailArgs {
    const __bf16* h;   // [B][H][W][32]
    const float* x;    // [B][H][W][2], the global skip
    float* y;          // [B][H][W][2]
    const bf16x8* w;   // tail fragments (pack_weights)
    const float* b;    // [2]
    int B, H, W;
};

__global__ __launch_bounds__(256, 2) void tail_fwd(TailArgs a) {
    constexpr int KSTEPS = ksteps_c(32);
    __shared__ __attribute__((aligned(16))) __bf16 in_t[(TL_NPRE * 256 / 4) * RB_PS];
    __shared__ bf16x8 w_s[KSTEPS * 64];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int i = threadIdx.x; i < KSTEPS * 64; i += 256) w_s[i] = a.w[i];
    const float bias0 = a.b[0], bias1 = a.b[1];
    const int tiles_x = (a.W + RB_TW - 1) / RB_TW, tiles_y = (a.H + RB_TH - 1) / RB_TH;
    const int ntiles = tiles_x * tiles_y * a.B;
    const unsigned img_bytes = (unsigned)a.H * (unsigned)a.W * 64u;  // < 2^31 (checked by the launcher)
    const unsigned xy_bytes = img_bytes / 8;
    bf16x8 pre[TL_NPRE];
    auto fetch = [&](int tile) {
        const int tx = tile % tiles_x, ty = (tile / tiles_x) % tiles_y, b = tile / (tiles_x * tiles_y);
        const int x0 = tx * RB_TW, y0 = ty * RB_TH;
        const __amdgpu_buffer_rsrc_t rs = oph_buf_rsrc(a.h + (size_t)b * a.H * a.W * 32, img_bytes);
#pragma unroll
        for (int n = 0; n < TL_NPRE; ++n) {
            const int i = threadIdx.x + 256 * n;
            const int pix = i >> 2, q = i & 3;
            const int yy = y0 + pix / RB_MW - 1, xx = x0 + pix % RB_MW - 1;
            const bool ok = i < RB_MH * RB_MW * 4 && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W;
            const unsigned off = ok ? (unsigned)((yy * a.W + xx) * 64 + 16 * q) : OPH_BUF_OOB;
            pre[n] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
        }
    };
    // outputs of the previous tile, stored after the next staging (as in
    // resblock_fwd: the prefetch stays the newest memory operation)
    u32x2 pv[2] = {{0u, 0u}, {0u, 0u}};
    unsigned pdst[2] = {OPH_BUF_OOB, OPH_BUF_OOB};
    __amdgpu_buffer_rsrc_t py = oph_buf_rsrc(a.y, xy_bytes);
    if ((int)blockIdx.x < ntiles) fetch(blockIdx.x);
    int gstep = (int)gridDim.x;
    asm volatile("" : "+s"(gstep));
    for (int tile = blockIdx.x; tile < ntiles; tile += gstep) {
        const int tx = tile % tiles_x, ty = (tile / tiles_x) % tiles_y, b = tile / (tiles_x * tiles_y);
        const int x0 = tx * RB_TW, y0 = ty * RB_TH;
        const size_t img = (size_t)b * a.H * a.W;
        int lane_o = lane, wave_o = wave;
        asm volatile("" : "+v"(lane_o), "+s"(wave_o));
        const int px = lane_o & 31, kh = lane_o >> 5;
        __syncthreads();  // the previous tile's LDS reads are done
#pragma unroll
        for (int n = 0; n < TL_NPRE; ++n) {
            const int i = threadIdx.x + 256 * n;
            *reinterpret_cast<bf16x8*>(in_t + (i >> 2) * RB_PS + 8 * (i & 3)) = pre[n];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 2; ++j) __builtin_amdgcn_raw_buffer_store_b64(pv[j], py, pdst[j], 0, 0);
        // the skip input of this tile's two rows (lane half 0: channels 0, 1)
        const int r0 = 2 * wave_o, x = x0 + px;
        const __amdgpu_buffer_rsrc_t rx = oph_buf_rsrc(a.x + img * 2, xy_bytes);
        unsigned dst[2];
        u32x2 xv[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int y = y0 + r0 + j;
            dst[j] = kh == 0 && y < a.H && x < a.W ? (unsigned)((y * a.W + x) * 8) : OPH_BUF_OOB;
            xv[j] = __builtin_bit_cast(u32x2, __builtin_amdgcn_raw_buffer_load_b64(rx, dst[j], 0, 0));
        }
        if (tile + gstep < ntiles) fetch(tile + gstep);

        f32x16 acc[2] = {(f32x16)0.f, (f32x16)0.f};
        bf16x8 Xr[4][6], Wr[RB_C2_AHEAD + 1];
        auto none = []() {};
        rb_c2_pipe<0, 0>(in_t + (r0 * RB_MW + px) * RB_PS + 8 * kh, w_s + lane_o, Xr, Wr, acc, none);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const f32x2 xs = __builtin_bit_cast(f32x2, xv[j]);
            const f32x2 o = {(acc[j][0] + bias0) + xs[0], (acc[j][1] + bias1) + xs[1]};
            pv[j] = __builtin_bit_cast(u32x2, o);
            pdst[j] = dst[j];
        }
        py = oph_buf_rsrc(a.y + img * 2, xy_bytes);
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) __builtin_amdgcn_raw_buffer_store_b64(pv[j], py, pdst[j], 0, 0);
}

// ---- host helpers ------------------------------------------------------------------

template <int CIN, int COUT, typename TI = float, typename TO = float>
int conv_fwd(oph_ctx* ctx, const TI* in, const bf16x8* wf, const float* bias,
             const typename std::remove_reference<TO>::type* res, TO* out, int B, int H, int W, int relu) {
    ConvArgsT<TI, TO> a{in, wf, bias, res, out, B, H, W, relu};
    const dim3 grid((unsigned)((W + CT_W - 1) / CT_W), (unsigned)((H + CT_H - 1) / CT_H), (unsigned)B);
    if (B > 65535 || grid.y > 65535) return oph_fail(ctx, OPH_EINVAL, "convolution batch or height too large");
    OPH_LAUNCH(ctx, "conv3x3_fwd", (conv3x3_fwd<CIN, COUT, TI, TO>), grid, dim3(64 * CT_WAVES), 0, a);
    return OPH_OK;
}

template <int CIN, int COUT>
int conv_dw(oph_ctx* ctx, oph_cnn* net, const float* dy, const float* x, int B, int H, int W, float* gw,
            float* gb) {
    constexpr int NCOL = 9 * CIN + 1;
    DwArgs a{dy, x, net->part, B, H, W, (W + CT_W - 1) / CT_W, (H + CT_H - 1) / CT_H};
    const int64_t ntiles = (int64_t)a.tiles_x * a.tiles_y * B;
    const int groups = (int)std::min<int64_t>(ntiles, DW_MAX_GROUPS);
    const size_t need = (size_t)groups * 32 * NCOL * sizeof(float);
    if (need > net->part_bytes) return oph_fail(ctx, OPH_ENOMEM, "weight-gradient workspace too small");
    if (ntiles > INT32_MAX) return oph_fail(ctx, OPH_EINVAL, "too many pixels");
    OPH_LAUNCH(ctx, "conv3x3_dw", (conv3x3_dw<CIN, COUT>), dim3((unsigned)groups), dim3(64 * DW_WAVES), 0, a);
    OPH_LAUNCH(ctx, "dw_reduce", dw_reduce, dim3((COUT * NCOL + 3) / 4), dim3(256), 0, net->part, groups, COUT,
               NCOL, gw, gb);
    return OPH_OK;
}

// dX = conv(dY, flipped/transposed W): input channels COUT, output CIN
template <int CIN, int COUT>
int conv_dx(oph_ctx* ctx, const float* dy, const bf16x8* wfrag, float* dx, const float* res, int B, int H, int W) {
    return conv_fwd<COUT, CIN>(ctx, dy, wfrag, nullptr, res, dx, B, H, W, 0);
}

constexpr int ksteps(int cin) { return (9 * cin + 15) / 16; }

// fragment offsets: forward (head, block b conv1/conv2, tail) and input-
// gradient (tail, block b conv2/conv1) convolutions, 64 fragments per k step
struct Frags {
    int NB;
    int64_t head() const { return 0; }
    int64_t conv1(int b) const { return 64LL * (ksteps(2) + (int64_t)b * 2 * ksteps(32)); }
    int64_t conv2(int b) const { return conv1(b) + 64LL * ksteps(32); }
    int64_t tail() const { return conv1(NB); }
    int64_t total() const { return tail() + 64LL * ksteps(32); }
    // input-gradient convolutions
    int64_t tail_t() const { return 0; }
    int64_t conv2_t(int b) const { return 64LL * (ksteps(2) + (int64_t)b * 2 * ksteps(32)); }
    int64_t conv1_t(int b) const { return conv2_t(b) + 64LL * ksteps(32); }
};

int pack(oph_ctx* ctx, oph_cnn* net, bool transposed) {
    const Layout L{net->ch, net->blocks};
    const Frags F{net->blocks};
    PackPlan plan{};
    int n = 0;
    if (!transposed) {
        plan.l[n++] = {L.head_w(), F.head(), 2, 32, 0};
        for (int b = 0; b < net->blocks; ++b) {
            plan.l[n++] = {L.w1(b), F.conv1(b), 32, 32, 0};
            plan.l[n++] = {L.w2(b), F.conv2(b), 32, 32, 0};
        }
        plan.l[n++] = {L.tail_w(), F.tail(), 32, 2, 0};
    } else {
        plan.l[n++] = {L.tail_w(), F.tail_t(), 2, 32, 1};
        for (int b = 0; b < net->blocks; ++b) {
            plan.l[n++] = {L.w2(b), F.conv2_t(b), 32, 32, 1};
            plan.l[n++] = {L.w1(b), F.conv1_t(b), 32, 32, 1};
        }
    }
    OPH_LAUNCH(ctx, "pack_weights", pack_weights, dim3((unsigned)((64 * ksteps(32) + 255) / 256), (unsigned)n),
               dim3(256), 0, net->params, transposed ? net->wb : net->wf, plan);
    if (!transposed) net->wf_rev = net->params_rev;
    return OPH_OK;
}

uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// activation workspace: x copy handled by caller; h_0..h_NB and a_1..a_NB,
// gradients dh, da, and the output
struct Acts {
    float* h[CMAX + 1];
    float* a[CMAX + 1];
    float* y;
    float* g0;
    float* g1;
    float* g2;
};

int ensure_workspace(oph_ctx* ctx, oph_cnn* net, int64_t npix, Acts& A) {
    const size_t per = (size_t)npix * net->ch * sizeof(float);
    const size_t need = per * (2 * net->blocks + 1 + 3) + (size_t)npix * 2 * sizeof(float);
    if (need > net->act_bytes) {
        if (net->act) hipFree(net->act);
        net->act = nullptr;
        net->act_bytes = 0;
        if (hipMalloc(&net->act, need) != hipSuccess) {
            hipGetLastError();
            return oph_fail(ctx, OPH_ENOMEM, "CNN activation workspace (%zu bytes)", need);
        }
        net->act_bytes = need;
    }
    char* p = reinterpret_cast<char*>(net->act);
    for (int b = 0; b <= net->blocks; ++b) { A.h[b] = reinterpret_cast<float*>(p); p += per; }
    for (int b = 1; b <= net->blocks; ++b) { A.a[b] = reinterpret_cast<float*>(p); p += per; }
    A.g0 = reinterpret_cast<float*>(p); p += per;
    A.g1 = reinterpret_cast<float*>(p); p += per;
    A.g2 = reinterpret_cast<float*>(p); p += per;
    A.y = reinterpret_cast<float*>(p);
    const size_t pneed = (size_t)DW_MAX_GROUPS * 32 * (9 * 32 + 1) * sizeof(float);
    if (pneed > net->part_bytes) {
        if (net->part) hipFree(net->part);
        net->part = nullptr;
        net->part_bytes = 0;
        if (hipMalloc(&net->part, pneed) != hipSuccess) {
            hipGetLastError();
            return oph_fail(ctx, OPH_ENOMEM, "CNN gradient workspace");
        }
        net->part_bytes = pneed;
    }
    return OPH_OK;
}

int forward(oph_ctx* ctx, oph_cnn* net, const float* x, int B, int H, int W, Acts& A) {
    const Layout L{net->ch, net->blocks};
    const float* P = net->params;
    const Frags F{net->blocks};
    const bf16x8* WF = net->wf;
    int rc;
    if (net->wf_rev != net->params_rev && (rc = pack(ctx, net, false))) return rc;
    if ((rc = conv_fwd<2, 32>(ctx, x, WF + F.head(), P + L.head_b(), nullptr, A.h[0], B, H, W, 0))) return rc;
    for (int b = 0; b < net->blocks; ++b) {
        if ((rc = conv_fwd<32, 32>(ctx, A.h[b], WF + F.conv1(b), P + L.b1(b), nullptr, A.a[b + 1], B, H, W, 1)))
            return rc;
        if ((rc = conv_fwd<32, 32>(ctx, A.a[b + 1], WF + F.conv2(b), P + L.b2(b), A.h[b], A.h[b + 1], B, H, W, 0)))
            return rc;
    }
    return conv_fwd<32, 2>(ctx, A.h[net->blocks], WF + F.tail(), P + L.tail_b(), x, A.y, B, H, W, 0);
}

// Inference forward with BF16 activations between layers: head conv, the
// fused residual blocks (ReLU outputs stay in LDS), tail conv; x and y stay
// float32 and y is written straight to `y`.
int forward_bf16(oph_ctx* ctx, oph_cnn* net, const float* x, float* y, int B, int H, int W, Acts& A) {
    const Layout L{net->ch, net->blocks};
    const float* P = net->params;
    const Frags F{net->blocks};
    const bf16x8* WF = net->wf;
    int rc;
    if (net->wf_rev != net->params_rev && (rc = pack(ctx, net, false))) return rc;
    auto h = [&](int b) { return reinterpret_cast<__bf16*>(A.h[b]); };
    // with blocks, the first one computes the head on its staged tiles
    if (net->blocks == 0 &&
        (rc = conv_fwd<2, 32, float, __bf16>(ctx, x, WF + F.head(), P + L.head_b(), nullptr, h(0), B, H, W, 0)))
        return rc;
    // persistent: two workgroups per CU (LDS-bound), each walking tiles
    const int64_t ntiles = (int64_t)((W + RB_TW - 1) / RB_TW) * ((H + RB_TH - 1) / RB_TH) * B;
    if (ntiles > INT32_MAX) return oph_fail(ctx, OPH_EINVAL, "too many pixels");
    // the fused block addresses one image's activations as a range-checked
    // buffer with 32-bit offsets
    if ((int64_t)H * W * 64 >= (1LL << 31))
        return oph_fail(ctx, OPH_ENOTSUP, "image %dx%d too large for the BF16 inference path (< 2^25 pixels)", H, W);
    const unsigned grid = (unsigned)std::min<int64_t>(ntiles, 2LL * std::max(ctx->cu_count, 1));
    const unsigned pgrid = (unsigned)std::min<int64_t>(ntiles, std::max(ctx->cu_count, 1));  // one per CU
    for (int b = 0; b < net->blocks; ++b) {
        const ResArgs r{h(b), h(b + 1), WF + F.conv1(b), WF + F.conv2(b), P + L.b1(b), P + L.b2(b), B, H, W,
                        x, WF + F.head(), P + L.head_b()};
        if (b == 0)
            OPH_LAUNCH(ctx, "resblock_fwd", resblock_fwd<true>, dim3(grid), dim3(256), 0, r);
        else
            OPH_LAUNCH(ctx, "resblock_pipe", resblock_pipe, dim3(pgrid), dim3(512), 0, r);
    }
    // three workgroups per CU fit (49 KB of LDS, 142 registers)
    const TailArgs t{h(net->blocks), x, y, WF + F.tail(), P + L.tail_b(), B, H, W};
    const unsigned tgrid = (unsigned)std::min<int64_t>(ntiles, 3LL * std::max(ctx->cu_count, 1));
    OPH_LAUNCH(ctx, "tail_fwd", tail_fwd, dim3(tgrid), dim3(256), 0, t);
    return OPH_OK;
}

// gradients of loss = mean (y - t)^2 into net->grads; loss into *loss
int backward(oph_ctx* ctx, oph_cnn* net, const float* x, const float* t, int B, int H, int W, Acts& A,
             double* loss) {
    const Layout L{net->ch, net->blocks};
    const float* P = net->params;
    float* G = net->grads;
    const int64_t npix = (int64_t)B * H * W;
    const int64_t n2 = npix * 2, nc = npix * net->ch;
    const unsigned nb = (unsigned)((n2 + 255) / 256);
    float* dy = A.g2;  // 2-channel output gradient (fits in a C-channel buffer)
    OPH_LAUNCH(ctx, "mse_grad", mse_grad, dim3(nb), dim3(256), 0, A.y, t, dy, n2, net->red);
    OPH_LAUNCH(ctx, "sum_partials", sum_partials, dim3(1), dim3(64), 0, net->red, (int)nb);
    const Frags F{net->blocks};
    const bf16x8* WB = net->wb;
    int rc;
    if ((rc = pack(ctx, net, true))) return rc;
    // tail: y = x + conv(h_NB)
    if ((rc = conv_dw<32, 2>(ctx, net, dy, A.h[net->blocks], B, H, W, G + L.tail_w(), G + L.tail_b()))) return rc;
    float* dh = A.g0;
    if ((rc = conv_dx<32, 2>(ctx, dy, WB + F.tail_t(), dh, nullptr, B, H, W))) return rc;
    for (int b = net->blocks - 1; b >= 0; --b) {
        // h_{b+1} = h_b + conv2(a); a = relu(conv1(h_b))
        float* da = A.g1;
        if ((rc = conv_dw<32, 32>(ctx, net, dh, A.a[b + 1], B, H, W, G + L.w2(b), G + L.b2(b)))) return rc;
        if ((rc = conv_dx<32, 32>(ctx, dh, WB + F.conv2_t(b), da, nullptr, B, H, W))) return rc;
        OPH_LAUNCH(ctx, "relu_mask", relu_mask, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, da, A.a[b + 1], nc);
        if ((rc = conv_dw<32, 32>(ctx, net, da, A.h[b], B, H, W, G + L.w1(b), G + L.b1(b)))) return rc;
        // dh_b = dh_{b+1} + conv1^T(da)
        if ((rc = conv_dx<32, 32>(ctx, da, WB + F.conv1_t(b), A.g2, dh, B, H, W))) return rc;
        float* tmp = dh;
        dh = A.g2;
        A.g2 = tmp;
    }
    if ((rc = conv_dw<2, 32>(ctx, net, dh, x, B, H, W, G + L.head_w(), G + L.head_b()))) return rc;
    if (loss) {
        double s = 0.0;
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(&s, net->red + nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        *loss = s / (double)n2;
    }
    return OPH_OK;
}

}  // namespace

extern "C" {

int oph_cnn_create(oph_ctx* ctx, int blocks, int channels, uint64_t seed, oph_cnn** out) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!out) return oph_fail(ctx, OPH_EINVAL, "null output");
    if (channels != 32 || blocks < 0 || blocks > MAX_BLOCKS)
        return oph_fail(ctx, OPH_ENOTSUP, "channels must be 32 and blocks 0..%d", MAX_BLOCKS);
    oph_cnn* n = new oph_cnn();
    n->blocks = blocks;
    n->ch = channels;
    const Layout L{channels, blocks};
    n->nparams = L.total();
    const size_t pb = (size_t)n->nparams * sizeof(float);
    if (hipMalloc(&n->params, pb) != hipSuccess || hipMalloc(&n->grads, pb) != hipSuccess ||
        hipMalloc(&n->m, pb) != hipSuccess || hipMalloc(&n->v, pb) != hipSuccess ||
        hipMalloc(&n->wf, (size_t)Frags{blocks}.total() * sizeof(bf16x8)) != hipSuccess ||
        hipMalloc(&n->wb, (size_t)Frags{blocks}.total() * sizeof(bf16x8)) != hipSuccess ||
        hipMalloc(&n->red, (size_t)(1 << 22) * sizeof(double)) != hipSuccess) {
        hipGetLastError();
        oph_cnn_destroy(ctx, n);
        return oph_fail(ctx, OPH_ENOMEM, "CNN parameter allocation");
    }
    // He-uniform init (seeded SplitMix64); the residual branch's second
    // convolution and the tail start small so the network starts near identity
    std::vector<float> h((size_t)n->nparams, 0.f);
    uint64_t s = seed;
    auto fill = [&](int64_t off, int64_t cnt, int fan_in, float gain) {
        const float lim = gain * sqrtf(6.f / (float)fan_in);
        for (int64_t i = 0; i < cnt; ++i) {
            const double u = (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0);
            h[off + i] = (float)((2.0 * u - 1.0) * lim);
        }
    };
    fill(L.head_w(), (int64_t)channels * 18, 18, 1.f);
    for (int b = 0; b < blocks; ++b) {
        fill(L.w1(b), (int64_t)channels * 9 * channels, 9 * channels, 1.f);
        fill(L.w2(b), (int64_t)channels * 9 * channels, 9 * channels, 0.1f);
    }
    fill(L.tail_w(), 2 * 9 * (int64_t)channels, 9 * channels, 0.1f);
    if (hipMemcpy(n->params, h.data(), pb, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(n->m, 0, pb) != hipSuccess || hipMemset(n->v, 0, pb) != hipSuccess ||
        hipMemset(n->grads, 0, pb) != hipSuccess) {
        hipGetLastError();
        oph_cnn_destroy(ctx, n);
        return oph_fail(ctx, OPH_EDEVICE, "CNN parameter upload");
    }
    *out = n;
    return OPH_OK;
}

void oph_cnn_destroy(oph_ctx* ctx, oph_cnn* n) {
    if (!n) return;
    if (ctx) hipSetDevice(ctx->device);
    hipDeviceSynchronize();
    for (void* p : {(void*)n->params, (void*)n->grads, (void*)n->m, (void*)n->v, (void*)n->wf, (void*)n->wb, (void*)n->act,
                    (void*)n->part, (void*)n->red})
        if (p) hipFree(p);
    delete n;
}

int64_t oph_cnn_num_params(const oph_cnn* n) { return n ? n->nparams : 0; }

int oph_cnn_get_params(oph_ctx* ctx, oph_cnn* n, float* host, int which) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!n || !host) return oph_fail(ctx, OPH_EINVAL, "null argument");
    const float* src = which == 1 ? n->grads : n->params;
    OPH_HIP_CHECK(ctx, hipMemcpyAsync(host, src, (size_t)n->nparams * 4, hipMemcpyDeviceToHost, ctx->stream));
    OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return OPH_OK;
}

int oph_cnn_set_params(oph_ctx* ctx, oph_cnn* n, const float* host) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!n || !host) return oph_fail(ctx, OPH_EINVAL, "null argument");
    OPH_HIP_CHECK(ctx, hipMemcpyAsync(n->params, host, (size_t)n->nparams * 4, hipMemcpyHostToDevice, ctx->stream));
    OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    n->params_rev += 1;
    n->step = 0;
    OPH_HIP_CHECK(ctx, hipMemsetAsync(n->m, 0, (size_t)n->nparams * 4, ctx->stream));
    OPH_HIP_CHECK(ctx, hipMemsetAsync(n->v, 0, (size_t)n->nparams * 4, ctx->stream));
    OPH_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return OPH_OK;
}

int oph_cnn_forward(oph_ctx* ctx, oph_cnn* n, const float* x, float* y, int B, int H, int W, int flags) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!n || !x || !y || B < 1 || H < 1 || W < 1) return oph_fail(ctx, OPH_EINVAL, "bad argument");
    const int64_t npix = (int64_t)B * H * W;
    Acts A;
    if ((rc = ensure_workspace(ctx, n, npix, A))) return rc;
    const size_t bytes = (size_t)npix * 2 * sizeof(float);
    const float* dx = x;
    float* dyv = y;
    if (!(flags & OPH_DEVICE_PTR)) {
        float* s = static_cast<float*>(oph_scratch(ctx, SLOT_IN, 2 * bytes));
        if (!s) return OPH_ENOMEM;
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(s, x, bytes, hipMemcpyHostToDevice, ctx->stream));
        dx = s;
        dyv = s + npix * 2;
    }
    if (flags & OPH_CNN_BF16_ACTS) {
        if ((rc = forward_bf16(ctx, n, dx, dyv, B, H, W, A))) return rc;
    } else {
        if ((rc = forward(ctx, n, dx, B, H, W, A))) return rc;
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(dyv, A.y, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    }
    if (!(flags & OPH_DEVICE_PTR))
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(y, dyv, bytes, hipMemcpyDeviceToHost, ctx->stream));
    return oph_finish(ctx, flags);
}

static int train_or_grad(oph_ctx* ctx, oph_cnn* n, const float* x, const float* t, int B, int H, int W,
                         float lr, double* loss, int flags, bool update) {
    int rc = oph_enter(ctx);
    if (rc) return rc;
    if (!n || !x || !t || B < 1 || H < 1 || W < 1) return oph_fail(ctx, OPH_EINVAL, "bad argument");
    const int64_t npix = (int64_t)B * H * W;
    if ((npix * 2 + 255) / 256 >= (1 << 22)) return oph_fail(ctx, OPH_EINVAL, "batch too large");
    Acts A;
    if ((rc = ensure_workspace(ctx, n, npix, A))) return rc;
    const size_t bytes = (size_t)npix * 2 * sizeof(float);
    const float* dx = x;
    const float* dt = t;
    if (!(flags & OPH_DEVICE_PTR)) {
        float* s = static_cast<float*>(oph_scratch(ctx, SLOT_IN, 2 * bytes));
        if (!s) return OPH_ENOMEM;
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(s, x, bytes, hipMemcpyHostToDevice, ctx->stream));
        OPH_HIP_CHECK(ctx, hipMemcpyAsync(s + npix * 2, t, bytes, hipMemcpyHostToDevice, ctx->stream));
        dx = s;
        dt = s + npix * 2;
    }
    if ((rc = forward(ctx, n, dx, B, H, W, A))) return rc;
    if ((rc = backward(ctx, n, dx, dt, B, H, W, A, loss))) return rc;
    if (update) {
        n->params_rev += 1;
        n->step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - powf(b1, (float)n->step), c2 = 1.f - powf(b2, (float)n->step);
        OPH_LAUNCH(ctx, "adam_step", adam_step, dim3((unsigned)((n->nparams + 255) / 256)), dim3(256), 0, n->params,
                   n->grads, n->m, n->v, n->nparams, lr, b1, b2, 1e-8f, c1, c2);
    }
    return oph_finish(ctx, flags);
}

int oph_cnn_train_step(oph_ctx* ctx, oph_cnn* n, const float* x, const float* t, int B, int H, int W, float lr,
                       double* loss, int flags) {
    return train_or_grad(ctx, n, x, t, B, H, W, lr, loss, flags, true);
}

int oph_cnn_gradients(oph_ctx* ctx, oph_cnn* n, const float* x, const float* t, int B, int H, int W, double* loss,
                      int flags) {
    return train_or_grad(ctx, n, x, t, B, H, W, 0.f, loss, flags, false);
}

// file: "OPHCNN1\0", int32 blocks, int32 channels, int64 nparams, float32 params
int oph_cnn_save(oph_ctx* ctx, oph_cnn* n, const char* path) {
    if (!n || !path) return oph_fail(ctx, OPH_EINVAL, "null argument");
    std::vector<float> h((size_t)n->nparams);
    int rc = oph_cnn_get_params(ctx, n, h.data(), 0);
    if (rc) return rc;
    FILE* f = fopen(path, "wb");
    if (!f) return oph_fail(ctx, OPH_EIO, "cannot write %s", path);
    const char magic[8] = {'O', 'P', 'H', 'C', 'N', 'N', '1', 0};
    const int32_t hdr[2] = {n->blocks, n->ch};
    const int64_t np = n->nparams;
    bool ok = fwrite(magic, 1, 8, f) == 8 && fwrite(hdr, 4, 2, f) == 2 && fwrite(&np, 8, 1, f) == 1 &&
              fwrite(h.data(), 4, h.size(), f) == h.size();
    fclose(f);
    return ok ? OPH_OK : oph_fail(ctx, OPH_EIO, "short write to %s", path);
}

int oph_cnn_load(oph_ctx* ctx, oph_cnn* n, const char* path) {
    if (!n || !path) return oph_fail(ctx, OPH_EINVAL, "null argument");
    FILE* f = fopen(path, "rb");
    if (!f) return oph_fail(ctx, OPH_EIO, "cannot read %s", path);
    char magic[8];
    int32_t hdr[2];
    int64_t np = 0;
    bool ok = fread(magic, 1, 8, f) == 8 && memcmp(magic, "OPHCNN1", 8) == 0 && fread(hdr, 4, 2, f) == 2 &&
              fread(&np, 8, 1, f) == 1 && hdr[0] == n->blocks && hdr[1] == n->ch && np == n->nparams;
    std::vector<float> h(ok ? (size_t)np : 0);
    ok = ok && fread(h.data(), 4, h.size(), f) == h.size();
    fclose(f);
    if (!ok) return oph_fail(ctx, OPH_EIO, "%s is not a matching network file", path);
    return oph_cnn_set_params(ctx, n, h.data());
}

}  // extern "C"
